! Drives libnegf_transport.so from Fortran through the ISO_C_BINDING module,
! the way phinfortran's driver would (negf/phinfortran.f90:1196-1251), on a
! monatomic chain whose answers are known in closed form:
!   surface GF  g = (w - sqrt(w^2 - 4t^2)) / (2 t^2) with w = E - 2k/m, t = -k/m
!   T(E) = 1 inside the band 0 < E < 4k/m, 0 outside, and T = 0 at ie = 1.
program fortran_abi_check
    use iso_c_binding
    use negf_transport
    implicit none
    type(c_ptr) :: ctx
    type(nt_params) :: p
    integer(c_int) :: st
    integer, parameter :: L = 4, ne = 9
    complex(c_double_complex) :: Haml(1,1), Hamlv(1,1), Hamr(1,1), Hamrv(1,1)
    complex(c_double_complex) :: hamd(L,1,1), hamdv(L,1,1), hamld(1,1), hamrd(1,1)
    complex(c_double_complex) :: gl_adv(1,1), gr_ret(1,1), gl_plus(1,1), gr_minus(1,1)
    complex(c_double_complex) :: gml_adv(1,1), gmr_ret(1,1), sl_adv(1,1), sr_ret(1,1)
    complex(c_double_complex) :: G_D_RL(1,1), G_D_LR(1,1), G_L(1,1), G_R(1,1)
    complex(c_double_complex) :: g_exact, w
    real(c_double) :: E, transl, transr, eta, Egrid(ne), tl(ne), tr(ne)
    integer :: i, nbad

    st = nt_create(ctx, 0_c_int)
    if (st /= 0) stop 'nt_create failed'
    call nt_params_default(p)
    p%eta = 1.0d-9
    p%eta0 = 1.0d-10
    p%convergence = 1.0d-14
    p%max_iter = 500
    ! chain: spring k = 1, mass m = 1 -> D00 = 2, V = -1
    Haml = (2.0d0, 0.0d0); Hamr = Haml
    Hamlv = (-1.0d0, 0.0d0); Hamrv = Hamlv
    hamd = (2.0d0, 0.0d0); hamdv = (-1.0d0, 0.0d0)
    hamld = (-1.0d0, 0.0d0); hamrd = (-1.0d0, 0.0d0)
    nbad = 0

    E = 1.3d0
    call nt_surface_green_tdb_f(ctx, p, E, Haml, Hamr, Hamlv, Hamrv, gl_adv, gr_ret, gl_plus, gr_minus, &
                                gml_adv, gmr_ret, sl_adv, sr_ret)
    eta = E * p%eta + p%eta0
    w = cmplx(E - 2.0d0, eta, kind=c_double)
    g_exact = (w - sqrt(w * w - 4.0d0)) / 2.0d0
    if (abs(aimag(g_exact)) > 0 .and. aimag(g_exact) > 0) g_exact = (w + sqrt(w * w - 4.0d0)) / 2.0d0
    print '(a, 2es24.15)', 'gr_ret      ', gr_ret(1,1)
    print '(a, 2es24.15)', 'closed_form ', g_exact
    if (abs(gr_ret(1,1) - g_exact) > 1.0d-9) nbad = nbad + 1
    if (abs(gl_adv(1,1) - conjg(g_exact)) > 1.0d-9) nbad = nbad + 1

    ! get_G_full_tdb_mpi passes the retarded left quantities
    call nt_get_G_full_tdb_f(ctx, p, .false., E, transpose(conjg(gl_adv)), gr_ret, -gml_adv, gmr_ret, &
                             transpose(conjg(sl_adv)), sr_ret, transl, transr, G_D_RL, G_D_LR, G_L, G_R, &
                             hamd, hamdv, hamld, hamrd, Haml, Hamr)
    print '(a, 2f20.14)', 'T_LR T_RL   ', transl, transr
    if (abs(transl - 1.0d0) > 1.0d-6 .or. abs(transr - 1.0d0) > 1.0d-6) nbad = nbad + 1

    ! the ie loop of one k-point in one call; first point zeroed like ie = 1
    do i = 1, ne
        Egrid(i) = (dble(i - 1) / dble(ne - 1) * 2.2d0)**2
    end do
    Egrid(1) = 0.001d0**2
    call nt_solve_tdb_f(ctx, p, Egrid, Haml, Hamlv, Hamr, Hamrv, hamd, hamdv, hamld, hamrd, tl, tr)
    do i = 1, ne
        print '(a, i3, 3f20.14)', 'solve ', i, Egrid(i), tl(i), tr(i)
        if (i == 1) then
            if (tl(i) /= 0.0d0 .or. tr(i) /= 0.0d0) nbad = nbad + 1
        else if (Egrid(i) < 3.9d0) then
            if (abs(tl(i) - 1.0d0) > 1.0d-5) nbad = nbad + 1
        else if (Egrid(i) > 4.1d0) then
            if (tl(i) > 1.0d-6) nbad = nbad + 1
        end if
    end do
    st = nt_destroy(ctx)
    if (nbad == 0) then
        print '(a)', 'FORTRAN_ABI_OK'
    else
        print '(a, i4)', 'FORTRAN_ABI_FAIL ', nbad
    end if
end program fortran_abi_check
