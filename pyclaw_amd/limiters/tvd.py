"""Wave-limiter ids understood by the classic kernels (classic/philim.f90:4-57).

Ids outside 1..5 behave as minmod, as the reference's computed GOTO falls
through to the minmod branch; 0 means no limiting for that wave family.
"""
minmod = 1
superbee = 2
vanleer = 3
MC = 4
beam_warming = 5
