! negf_transport.f90 -- ISO_C_BINDING interface to libnegf_transport.so.
!
! Gives the Fortran driver of phinfortran a drop-in for its per-(omega,k)
! solve.  The bind(C) interfaces map one-to-one onto include/negf_transport.h;
! the nt_* Fortran wrappers below keep the reference's argument order and
! meaning and its error behaviour (stop 'Matrix is numerically singular!',
! negf/util.f90:44-58) so a call site changes only in its name:
!
!   call surface_green_tdb(mm, gl_adv, ...)       negf/surface.f90:385
!     -> call nt_surface_green_tdb_f(ctx, p, E, Haml, Hamr, Hamlv, Hamrv, gl_adv, ...)
!   call get_G_full_tdb(ie, mm, gl, gr, ...)       negf/lesser.f90:475
!     -> call nt_get_G_full_tdb_f(ctx, p, ie == 1, E, gl, gr, ..., transl, transr, ...)
!   the whole ie loop of one k-point                negf/phinfortran.f90:1196-1251
!     -> call nt_solve_tdb_f(ctx, p, Egrid(:, nn), Haml, ..., transl(:, nn), transr(:, nn))
module negf_transport
    use iso_c_binding
    implicit none

    integer(c_int), parameter :: NT_OK = 0, NT_ERR_NAN = -1, NT_ERR_NOT_CONVERGED = -2, &
                                 NT_ERR_ARG = -3, NT_ERR_HIP = -4, NT_ERR_NOMEM = -5
    integer(c_int), parameter :: NT_ETA_RELATIVE = 0, NT_ETA_ABSOLUTE = 1
    integer(c_int), parameter :: NT_PREC_FP64 = 0, NT_PREC_MIXED = 1
    integer(c_int), parameter :: NT_DOS_FROM_ETA = 0, NT_DOS_NEGF = 1, NT_DOS_TB = 2

    type, bind(C) :: nt_params
        real(c_double) :: eta, eta0, convergence, device_eta_factor, lambda_eta
        integer(c_int) :: eta_mode, max_iter, precision, circular, n_bloch_x, n_bloch_y, path_mode
        integer(c_int) :: dos_mode
    end type nt_params

    type, bind(C) :: nt_solve_extras
        type(c_ptr) :: gl_adv = c_null_ptr, gr_ret = c_null_ptr, gl_plus = c_null_ptr, gr_minus = c_null_ptr
        type(c_ptr) :: gml_adv = c_null_ptr, gmr_ret = c_null_ptr, sl_adv = c_null_ptr, sr_ret = c_null_ptr
        type(c_ptr) :: G_D_LR = c_null_ptr, G_D_RL = c_null_ptr, G_L = c_null_ptr, G_R = c_null_ptr
        type(c_ptr) :: sdos_l = c_null_ptr, sdos_r = c_null_ptr
        type(c_ptr) :: ldos = c_null_ptr, ddos = c_null_ptr
    end type nt_solve_extras

    type, bind(C) :: nt_modes_out
        type(c_ptr) :: trl = c_null_ptr, tlr = c_null_ptr, rrr = c_null_ptr, rll = c_null_ptr
        type(c_ptr) :: k_z_L_adv = c_null_ptr, k_z_L_ret = c_null_ptr, k_z_R_ret = c_null_ptr, k_z_R_adv = c_null_ptr
        type(c_ptr) :: n_prop = c_null_ptr, t_mode_L = c_null_ptr, t_mode_R = c_null_ptr, sums = c_null_ptr
    end type nt_modes_out

    ! Brillouin-zone unfolding inputs (nt_unfold_desc): the reciprocal
    ! vectors row-major (pass reshape(transpose(recivec), [9]) of the
    ! reference's recivec(3,3)), device pointers to ksc (transpose(k_grid)),
    ! the orbital positions (transpose(pos_l(:,1:3))) and the cell-atom
    ! positions positions(L_list(pidxL(o)),:) of the first image's orbitals.
    type, bind(C) :: nt_unfold_desc
        integer(c_int) :: n_bloch_x = 1, n_bloch_y = 1
        real(c_double) :: recivec(9) = 0.0d0, reci_uc_l(9) = 0.0d0, reci_uc_r(9) = 0.0d0
        type(c_ptr) :: ksc = c_null_ptr, pos_l = c_null_ptr, pos_uc_l = c_null_ptr
        type(c_ptr) :: pos_r = c_null_ptr, pos_uc_r = c_null_ptr
    end type nt_unfold_desc

    ! Unfolding outputs (device pointers, any may be null): bz_idx_*, k_xyz_*
    ! (n x 3 per point, the reference's k_xyz_* layout), the eight split
    ! sums per point and the (total, specular) sums per incident mode.
    type, bind(C) :: nt_unfold_out
        type(c_ptr) :: bz_L_adv = c_null_ptr, bz_L_ret = c_null_ptr, bz_R_ret = c_null_ptr, bz_R_adv = c_null_ptr
        type(c_ptr) :: k_xyz_L_adv = c_null_ptr, k_xyz_L_ret = c_null_ptr
        type(c_ptr) :: k_xyz_R_ret = c_null_ptr, k_xyz_R_adv = c_null_ptr
        type(c_ptr) :: split = c_null_ptr
        type(c_ptr) :: spec_trl = c_null_ptr, spec_tlr = c_null_ptr, spec_rrr = c_null_ptr, spec_rll = c_null_ptr
    end type nt_unfold_out

    ! Buttiker probes (nt_buttiker_tdb_batched_dev): temperature points to the
    ! L probe temperatures in device memory
    type, bind(C) :: nt_probe_desc
        type(c_ptr) :: temperature = c_null_ptr
        real(c_double) :: t_left = 300.0d0, t_right = 300.0d0
        real(c_double) :: B_probe = 5.0d-20*(2.0d0*3.1415926d0*1.0d12), T_probe = 430.0d0
        integer(c_int) :: lifetime = 0
    end type nt_probe_desc

    ! per point p: flux(L), jacobian(L,L) (jacobian(i,j) = d flux_i / d T_j),
    ! flux_left, flux_right, a_sum(L), gn_sum(L); device memory
    type, bind(C) :: nt_probe_out
        type(c_ptr) :: flux = c_null_ptr, jacobian = c_null_ptr
        type(c_ptr) :: flux_left = c_null_ptr, flux_right = c_null_ptr
        type(c_ptr) :: a_sum = c_null_ptr, gn_sum = c_null_ptr
    end type nt_probe_out

    interface
        subroutine nt_params_default(p) bind(C, name="nt_params_default")
            import :: nt_params
            type(nt_params), intent(out) :: p
        end subroutine

        integer(c_int) function nt_params_validate(p, nl, nr) bind(C, name="nt_params_validate")
            import :: c_int, nt_params
            type(nt_params), intent(in) :: p
            integer(c_int), value :: nl, nr
        end function

        integer(c_int) function nt_create(ctx, device) bind(C, name="nt_create")
            import :: c_int, c_ptr
            type(c_ptr), intent(out) :: ctx
            integer(c_int), value :: device
        end function

        integer(c_int) function nt_destroy(ctx) bind(C, name="nt_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        integer(c_int) function nt_set_batch_limit(ctx, n) bind(C, name="nt_set_batch_limit")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: n
        end function

        integer(c_int) function nt_surface_green_tdb(ctx, p, nl, nr, E, Haml, Hamr, Hamlv, Hamrv, &
                gl_adv, gr_ret, gl_plus, gr_minus, gml_adv, gmr_ret, sl_adv, sr_ret, iters) &
                bind(C, name="nt_surface_green_tdb")
            import :: c_int, c_ptr, c_double, nt_params
            type(c_ptr), value :: ctx
            type(nt_params), intent(in) :: p
            integer(c_int), value :: nl, nr
            real(c_double), value :: E
            type(c_ptr), value :: Haml, Hamr, Hamlv, Hamrv
            type(c_ptr), value :: gl_adv, gr_ret, gl_plus, gr_minus, gml_adv, gmr_ret, sl_adv, sr_ret, iters
        end function

        integer(c_int) function nt_get_G_full_tdb(ctx, p, first_point, nl, nr, b, L, E, gl, gr, gml, gmr, &
                sl, sr, hamd, hamdv, hamld, hamrd, haml, hamr, transl, transr, G_D_RL, G_D_LR, G_L, G_R) &
                bind(C, name="nt_get_G_full_tdb")
            import :: c_int, c_ptr, c_double, nt_params
            type(c_ptr), value :: ctx
            type(nt_params), intent(in) :: p
            integer(c_int), value :: first_point, nl, nr, b, L
            real(c_double), value :: E
            type(c_ptr), value :: gl, gr, gml, gmr, sl, sr, hamd, hamdv, hamld, hamrd, haml, hamr
            real(c_double), intent(out) :: transl, transr
            type(c_ptr), value :: G_D_RL, G_D_LR, G_L, G_R
        end function

        ! G_block (negf/device.f90:289-560, tb/device.f90:291-575): layer_list is an
        ! int array, Hd/Hdv/Hdv_back arrays of c_ptr (c_loc of each block copy)
        integer(c_int) function nt_G_block(ctx, p, nl, nr, L, layer_list, E, gl, gr, gml, gmr, sl, sr, &
                Hd, Hdv, Hdv_back, Hld, Hrd, Hrd_fwd, Haml, Hamr, transl, transr, G_D_RL, G_D_LR, G_L, G_R) &
                bind(C, name="nt_G_block")
            import :: c_int, c_ptr, c_double, nt_params
            type(c_ptr), value :: ctx
            type(nt_params), intent(in) :: p
            integer(c_int), value :: nl, nr, L
            type(c_ptr), value :: layer_list
            real(c_double), value :: E
            type(c_ptr), value :: gl, gr, gml, gmr, sl, sr, Hd, Hdv, Hdv_back, Hld, Hrd, Hrd_fwd, Haml, Hamr
            real(c_double), intent(out) :: transl, transr
            type(c_ptr), value :: G_D_RL, G_D_LR, G_L, G_R
        end function

        ! get_ldos (tb/device.f90:610-855); ldos(L, maxlen) column-major
        integer(c_int) function nt_get_ldos(ctx, p, nl, nr, L, layer_list, E, gl, gr, Hd, Hdv, Hdv_back, &
                Hld, Hrd, Hrd_fwd, maxlen, ldos, ddos) bind(C, name="nt_get_ldos")
            import :: c_int, c_ptr, c_double, nt_params
            type(c_ptr), value :: ctx
            type(nt_params), intent(in) :: p
            integer(c_int), value :: nl, nr, L, maxlen
            type(c_ptr), value :: layer_list
            real(c_double), value :: E
            type(c_ptr), value :: gl, gr, Hd, Hdv, Hdv_back, Hld, Hrd, Hrd_fwd, ldos
            real(c_double), intent(out) :: ddos
        end function

        ! vel_mat_tdb (negf/velocity_mat.f90:1692-3150) on device memory
        integer(c_int) function nt_modes_tdb_batched_dev(ctx, p, tb, nl, nr, npts, E, kidx, nk, Haml, Hamlv, &
                Hamr, Hamrv, gl_adv, gr_ret, gl_plus, gr_minus, gml_adv, gmr_ret, G_D_RL, G_D_LR, G_L, G_R, &
                a_z_L, a_z_R, first_point_mask, out, status) bind(C, name="nt_modes_tdb_batched_dev")
            import :: c_int, c_ptr, c_double, nt_params, nt_modes_out
            type(c_ptr), value :: ctx
            type(nt_params), intent(in) :: p
            integer(c_int), value :: tb, nl, nr, npts, nk
            type(c_ptr), value :: E, kidx, Haml, Hamlv, Hamr, Hamrv, gl_adv, gr_ret, gl_plus, gr_minus
            type(c_ptr), value :: gml_adv, gmr_ret, G_D_RL, G_D_LR, G_L, G_R
            real(c_double), value :: a_z_L, a_z_R
            type(c_ptr), value :: first_point_mask
            type(nt_modes_out), intent(in) :: out
            type(c_ptr), value :: status
        end function

        ! find_k_pc (negf/dyn.f90:770-889) for nmodes modes; reci_uc / recivec_pc
        ! row-major (reshape(transpose(reci_uc_l), [9])); other pointers device memory
        integer(c_int) function nt_unfold_pc_dev(ctx, nmodes, E, k_xyz, n_bloch_max, reci_uc, recivec_pc, n_pc, &
                nseg, seg_off, seg_rc, nent, coef, dr, kp, v, ncand) bind(C, name="nt_unfold_pc_dev")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nmodes, n_bloch_max, n_pc, nseg, nent
            type(c_ptr), value :: E, k_xyz
            real(c_double), intent(in) :: reci_uc(9), recivec_pc(9)
            type(c_ptr), value :: seg_off, seg_rc, coef, dr, kp, v, ncand
        end function

        ! find_k_pc_l / find_k_pc_r of the tight-binding run (tb/hamiltonian.f90:666-916):
        ! the primitive cell's H(k) from an element-wise table (el_off, coef, dr)
        integer(c_int) function nt_unfold_pc_tb_dev(ctx, nmodes, E, k_xyz, n_bloch_max, reci_uc, recivec_pc, &
                n_pc, el_off, nent, coef, dr, kp, v, ncand) bind(C, name="nt_unfold_pc_tb_dev")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nmodes, n_bloch_max, n_pc, nent
            type(c_ptr), value :: E, k_xyz
            real(c_double), intent(in) :: reci_uc(9), recivec_pc(9)
            type(c_ptr), value :: el_off, coef, dr, kp, v, ncand
        end function

        ! vel_mat_tdb with the Brillouin-zone unfolding (velocity_mat.f90:2020-2130, :3032-3099)
        integer(c_int) function nt_modes_unfold_tdb_batched_dev(ctx, p, tb, nl, nr, npts, E, kidx, nk, Haml, &
                Hamlv, Hamr, Hamrv, gl_adv, gr_ret, gl_plus, gr_minus, gml_adv, gmr_ret, G_D_RL, G_D_LR, G_L, &
                G_R, a_z_L, a_z_R, first_point_mask, out, uf, uo, status) &
                bind(C, name="nt_modes_unfold_tdb_batched_dev")
            import :: c_int, c_ptr, c_double, nt_params, nt_modes_out, nt_unfold_desc, nt_unfold_out
            type(c_ptr), value :: ctx
            type(nt_params), intent(in) :: p
            integer(c_int), value :: tb, nl, nr, npts, nk
            type(c_ptr), value :: E, kidx, Haml, Hamlv, Hamr, Hamrv, gl_adv, gr_ret, gl_plus, gr_minus
            type(c_ptr), value :: gml_adv, gmr_ret, G_D_RL, G_D_LR, G_L, G_R
            real(c_double), value :: a_z_L, a_z_R
            type(c_ptr), value :: first_point_mask
            type(nt_modes_out), intent(in) :: out
            type(nt_unfold_desc), intent(in) :: uf
            type(nt_unfold_out), intent(in) :: uo
            type(c_ptr), value :: status
        end function

        integer(c_int) function nt_buttiker_tdb_batched_dev(ctx, p, probe, nl, nr, b, L, npts, E, kidx, nk, &
                Haml, Hamlv, Hamr, Hamrv, Hamd, Hamdv, Hamld, Hamrd, first_point_mask, &
                T_LR, T_RL, status, iters, extras, out) bind(C, name="nt_buttiker_tdb_batched_dev")
            import :: c_int, c_ptr, nt_params, nt_probe_desc, nt_probe_out, nt_solve_extras
            type(c_ptr), value :: ctx
            type(nt_params), intent(in) :: p
            type(nt_probe_desc), intent(in) :: probe
            integer(c_int), value :: nl, nr, b, L, npts, nk
            type(c_ptr), value :: E, kidx, Haml, Hamlv, Hamr, Hamrv, Hamd, Hamdv, Hamld, Hamrd
            type(c_ptr), value :: first_point_mask, T_LR, T_RL, status, iters
            type(nt_solve_extras), intent(in) :: extras
            type(nt_probe_out), intent(in) :: out
        end function

        integer(c_int) function nt_solve_tdb_batched(ctx, p, nl, nr, b, L, npts, E, kidx, nk, &
                Haml, Hamlv, Hamr, Hamrv, Hamd, Hamdv, Hamld, Hamrd, first_point_mask, &
                T_LR, T_RL, status, iters, extras) bind(C, name="nt_solve_tdb_batched")
            import :: c_int, c_ptr, nt_params, nt_solve_extras
            type(c_ptr), value :: ctx
            type(nt_params), intent(in) :: p
            integer(c_int), value :: nl, nr, b, L, npts, nk
            type(c_ptr), value :: E, kidx, Haml, Hamlv, Hamr, Hamrv, Hamd, Hamdv, Hamld, Hamrd
            type(c_ptr), value :: first_point_mask, T_LR, T_RL, status, iters
            type(nt_solve_extras), intent(in) :: extras
        end function

        ! gen_ham (tb/hamiltonian.f90:444-519): H(k) of nk k-points from a
        ! phase-sum table of norb_i x norb_j sub-blocks; all pointers device memory
        integer(c_int) function nt_hk_assemble_orb_dev(ctx, nk, kcart, nseg, seg_off, seg_geom, nent, cstride, &
                coef, dr, rows, cols, out, ld, kstride, hermitize) bind(C, name="nt_hk_assemble_orb_dev")
            import :: c_int, c_ptr, c_long_long
            type(c_ptr), value :: ctx
            integer(c_int), value :: nk, nseg, nent, cstride, rows, cols, ld, hermitize
            integer(c_long_long), value :: kstride
            type(c_ptr), value :: kcart, seg_off, seg_geom, coef, dr, out
        end function

        function nt_status_string(status) bind(C, name="nt_status_string")
            import :: c_int, c_ptr
            integer(c_int), value :: status
            type(c_ptr) :: nt_status_string
        end function
    end interface

contains

    ! The reference stops the program on a singular block (negf/util.f90:44-58)
    ! and on NaN input (util.f90:29-32); the wrappers do the same.
    subroutine nt_check(status)
        integer(c_int), intent(in) :: status
        if (status > 0) then
            write(*,*) 'info'
            write(*,*) status
            stop 'Matrix is numerically singular!'
        else if (status == NT_ERR_NAN) then
            stop
        else if (status /= NT_OK) then
            write(*,*) 'negf_transport status', status
            stop 'negf_transport failed'
        end if
    end subroutine nt_check

    ! surface_green_tdb (negf/surface.f90:385-605) on the GPU.
    subroutine nt_surface_green_tdb_f(ctx, p, E, Haml, Hamr, Hamlv, Hamrv, &
                                      gl_adv, gr_ret, gl_plus, gr_minus, gml_adv, gmr_ret, sl_adv, sr_ret)
        type(c_ptr), intent(in) :: ctx
        type(nt_params), intent(in) :: p
        real(c_double), intent(in) :: E
        complex(c_double_complex), intent(in), target :: Haml(:,:), Hamr(:,:), Hamlv(:,:), Hamrv(:,:)
        complex(c_double_complex), intent(out), target :: gl_adv(:,:), gr_ret(:,:), gl_plus(:,:), gr_minus(:,:)
        complex(c_double_complex), intent(out), target :: gml_adv(:,:), gmr_ret(:,:), sl_adv(:,:), sr_ret(:,:)
        complex(c_double_complex), allocatable, target :: hl(:,:), hr(:,:), vl(:,:), vr(:,:)
        integer(c_int) :: st
        hl = Haml
        hr = Hamr
        vl = Hamlv
        vr = Hamrv
        st = nt_surface_green_tdb(ctx, p, int(size(Haml, 1), c_int), int(size(Hamr, 1), c_int), E, &
                c_loc(hl), c_loc(hr), c_loc(vl), c_loc(vr), c_loc(gl_adv), c_loc(gr_ret), &
                c_loc(gl_plus), c_loc(gr_minus), c_loc(gml_adv), c_loc(gmr_ret), c_loc(sl_adv), &
                c_loc(sr_ret), c_null_ptr)
        call nt_check(st)
    end subroutine nt_surface_green_tdb_f

    ! get_G_full_tdb (negf/lesser.f90:475-809) on the GPU.  hamd(L,b,b) and
    ! hamdv(L,b,b) are in the reference's layer-first layout and are regrouped
    ! layer by layer here.  gl, gml, sl are the retarded left quantities that
    ! get_G_full_tdb_mpi passes (transpose(dconjg(gl_adv)), -gml_adv, ...).
    subroutine nt_get_G_full_tdb_f(ctx, p, first_point, E, gl, gr, gml, gmr, sl, sr, &
                                   transl, transr, G_D_RL, G_D_LR, G_L, G_R, &
                                   hamd, hamdv, hamld, hamrd, haml, hamr)
        type(c_ptr), intent(in) :: ctx
        type(nt_params), intent(in) :: p
        logical, intent(in) :: first_point
        real(c_double), intent(in) :: E
        complex(c_double_complex), intent(in) :: gl(:,:), gr(:,:), gml(:,:), gmr(:,:), sl(:,:), sr(:,:)
        real(c_double), intent(out) :: transl, transr
        complex(c_double_complex), intent(out), target :: G_D_RL(:,:), G_D_LR(:,:), G_L(:,:), G_R(:,:)
        complex(c_double_complex), intent(in) :: hamd(:,:,:), hamdv(:,:,:)
        complex(c_double_complex), intent(in) :: hamld(:,:), hamrd(:,:), haml(:,:), hamr(:,:)
        complex(c_double_complex), allocatable, target :: d(:,:,:), dv(:,:,:)
        complex(c_double_complex), allocatable, target :: a1(:,:), a2(:,:), a3(:,:), a4(:,:), a5(:,:), a6(:,:)
        complex(c_double_complex), allocatable, target :: b1(:,:), b2(:,:), b3(:,:), b4(:,:)
        integer :: i, L, b
        integer(c_int) :: st, fp
        L = size(hamd, 1)
        b = size(hamd, 2)
        allocate(d(b, b, L), dv(b, b, max(L - 1, 1)))
        do i = 1, L
            d(:, :, i) = hamd(i, :, :)
        end do
        do i = 1, L - 1
            dv(:, :, i) = hamdv(i, :, :)
        end do
        a1 = gl; a2 = gr; a3 = gml; a4 = gmr; a5 = sl; a6 = sr
        b1 = hamld; b2 = hamrd; b3 = haml; b4 = hamr
        fp = 0
        if (first_point) fp = 1
        st = nt_get_G_full_tdb(ctx, p, fp, int(size(haml, 1), c_int), int(size(hamr, 1), c_int), int(b, c_int), &
                int(L, c_int), E, c_loc(a1), c_loc(a2), c_loc(a3), c_loc(a4), c_loc(a5), c_loc(a6), &
                c_loc(d), c_loc(dv), c_loc(b1), c_loc(b2), c_loc(b3), c_loc(b4), transl, transr, &
                c_loc(G_D_RL), c_loc(G_D_LR), c_loc(G_L), c_loc(G_R))
        call nt_check(st)
    end subroutine nt_get_G_full_tdb_f

    ! All energies of one k-point in one batched call: the body of the driver's
    ! ie loop (negf/phinfortran.f90:1196-1251) for surface GFs + RGF + Caroli.
    ! E(ne) is Egrid(:, nn); point 1 gets the reference's T = 0 (lesser.f90:802).
    subroutine nt_solve_tdb_f(ctx, p, E, Haml, Hamlv, Hamr, Hamrv, hamd, hamdv, hamld, hamrd, transl, transr)
        type(c_ptr), intent(in) :: ctx
        type(nt_params), intent(in) :: p
        real(c_double), intent(in), target :: E(:)
        complex(c_double_complex), intent(in) :: Haml(:,:), Hamlv(:,:), Hamr(:,:), Hamrv(:,:)
        complex(c_double_complex), intent(in) :: hamd(:,:,:), hamdv(:,:,:), hamld(:,:), hamrd(:,:)
        real(c_double), intent(out), target :: transl(:), transr(:)
        complex(c_double_complex), allocatable, target :: d(:,:,:), dv(:,:,:)
        complex(c_double_complex), allocatable, target :: hl(:,:), vl(:,:), hr(:,:), vr(:,:), ld(:,:), rd(:,:)
        integer(c_int), allocatable, target :: mask(:), status(:)
        real(c_double), allocatable, target :: ee(:)
        type(nt_solve_extras) :: ex
        integer :: i, L, b, ne
        integer(c_int) :: st
        L = size(hamd, 1)
        b = size(hamd, 2)
        ne = size(E)
        allocate(d(b, b, L), dv(b, b, max(L - 1, 1)), mask(ne), status(ne))
        do i = 1, L
            d(:, :, i) = hamd(i, :, :)
        end do
        do i = 1, L - 1
            dv(:, :, i) = hamdv(i, :, :)
        end do
        hl = Haml; vl = Hamlv; hr = Hamr; vr = Hamrv; ld = hamld; rd = hamrd; ee = E
        mask = 0
        mask(1) = 1
        st = nt_solve_tdb_batched(ctx, p, int(size(Haml, 1), c_int), int(size(Hamr, 1), c_int), int(b, c_int), &
                int(L, c_int), int(ne, c_int), c_loc(ee), c_null_ptr, 1_c_int, c_loc(hl), c_loc(vl), c_loc(hr), &
                c_loc(vr), c_loc(d), c_loc(dv), c_loc(ld), c_loc(rd), c_loc(mask), c_loc(transl), c_loc(transr), &
                c_loc(status), c_null_ptr, ex)
        call nt_check(st)
    end subroutine nt_solve_tdb_f

end module negf_transport
