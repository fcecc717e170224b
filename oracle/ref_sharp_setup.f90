! Harness for the reference SharpClaw module state (test infrastructure only).
! The reference sets these module variables from Python
! (src/pyclaw/sharpclaw/solver.py:302-328); this routine does the same from C.
! Configuration: component-wise WENO5 (lim_type=2, char_decomp=0), no capa,
! no tfluct solver, no f-wave.
subroutine ref_sharp_setup(meqn, mwaves, ng, maxnx, dxx, dyy) bind(c, name='ref_sharp_setup')
    use iso_c_binding
    use clawparams
    use workspace
    use reconstruct
    implicit none
    integer(c_int), value :: meqn, mwaves, ng, maxnx
    real(c_double), value :: dxx, dyy
    integer :: mw, nx

    if (allocated(dx)) call dealloc_clawparams()
    if (work_alloc) call dealloc_workspace(0.0)
    if (recon_alloc) call dealloc_recon_workspace(2, 0)
    num_dim = 2
    lim_type = 2
    weno_order = 5
    char_decomp = 0
    tfluct_solver = .false.
    fwave = .false.
    index_capa = 0
    num_waves = mwaves
    call alloc_clawparams()
    xlower = 0.d0
    xupper = 1.d0
    dx(1) = dxx
    dx(2) = dyy
    do mw = 1, mwaves
        mthlim(mw) = 1
    end do
    nx = maxnx + 2*ng
    call alloc_workspace(nx, ng, meqn, mwaves, 0)
    call alloc_recon_workspace(nx, ng, meqn, mwaves, 2, 0)
end subroutine ref_sharp_setup
