2
def2 TEST samples/arbitrary2qubitmatrixset/matrix10.gate
TEST 0 1
