
GRID      1       0.0       0.0       0.0  0
GRID      2       1.0       0.0       0.0  0
GRID      3       1.0       1.0       0.0  0
GRID      4       0.0       1.0       0.0  0
GRID      5       0.0       0.0       1.0  0
GRID      6       1.0       0.0       1.0  0
GRID      7       1.0       1.0       1.0  0
GRID      8       0.0       1.0       1.0  0

CTRIAX      1       11   1       2       3
CTRIAX      2       11   1       3       4
CTRIAX      3      11   5       6       7
CTRIAX      4       11   5       7       8
CTRIAX      5       1   1       2       5
CTRIAX      6       1   2       5       6
CTRIAX      7       1   2       3       7
CTRIAX      8       1   2       6       7
CTRIAX      9       1   3       4       8
CTRIAX      10       1   3       7       8
CTRIAX      11       1   1       4       8
CTRIAX      12       1   1       5       8
CTRIAX      13      -1   1       3       5
CTRIAX      14      -1   3       5       7
CTRIAX      15      -1   3       5       8
CTRIAX      16      -1   2       3       5
CTRIAX      17      -1   1       3       8
CTRIAX      18      -1   2       5       7


CTETRA      1       1   3       1       2       5
CTETRA      2       1   3       4       1       8
CTETRA      3       1   5       3       1       8
CTETRA      4       1   6       2       5       7
CTETRA      5       1   5       2       3       7
CTETRA      6       1   8       5       3       7
