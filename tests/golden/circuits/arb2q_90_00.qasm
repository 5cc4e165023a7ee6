2
def2 TEST samples/arbitrary2qubitmatrixset/matrix90.gate
TEST 0 1
