2
def2 TEST samples/arbitrary2qubitmatrixset/matrix80.gate
TEST 0 1
