2
def2 TEST samples/arbitrary2qubitmatrixset/matrix0.gate
TEST 0 1
