2
H 0
H 1
