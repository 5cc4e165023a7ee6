1
def1 HadamardTester circuits/hadtest.gate
HadamardTester 0
