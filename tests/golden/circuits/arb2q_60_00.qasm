2
def2 TEST samples/arbitrary2qubitmatrixset/matrix60.gate
TEST 0 1
