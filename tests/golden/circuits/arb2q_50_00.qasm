2
def2 TEST samples/arbitrary2qubitmatrixset/matrix50.gate
TEST 0 1
