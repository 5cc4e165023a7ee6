2
X 0
X 1
def2 TEST samples/arbitrary2qubitmatrixset/matrix40.gate
TEST 0 1
