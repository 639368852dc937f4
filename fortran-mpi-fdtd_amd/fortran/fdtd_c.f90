! fdtd_c.f90 -- iso_c_binding interface to include/fdtd.h for Fortran hosts.
!
! The derived types are bind(c) images of the C structs (field order and
! kinds match; Fortran names differ where C names differ only by case, e.g.
! `gam` for `gamma` next to `cp_gamma` for `Gamma`).  Handles are type(c_ptr).
module fdtd_c
    use iso_c_binding
    implicit none

    integer(c_int), parameter :: FDTD_OK = 0
    integer(c_int), parameter :: FDTD_FP64 = 0, FDTD_FP32 = 1
    integer(c_int), parameter :: FDTD_PML_NONE = 0, FDTD_PML_REF_Y = 1, FDTD_PML_ALL_FACES = 2
    integer(c_int), parameter :: FDTD_EX = 0, FDTD_EY = 1, FDTD_EZ = 2
    integer(c_int), parameter :: FDTD_HX = 3, FDTD_HY = 4, FDTD_HZ = 5
    integer(c_int), parameter :: FDTD_DX = 6, FDTD_DY = 7, FDTD_DZ = 8
    integer(c_int), parameter :: FDTD_REGION_BOX = 0, FDTD_REGION_SPHERE = 1, FDTD_REGION_CYLINDER = 2
    integer(c_int), parameter :: FDTD_GRAPHENE_CP4 = 0, FDTD_GRAPHENE_KUBO = 1
    integer(c_int), parameter :: FDTD_ENGINE_AUTO = 0, FDTD_ENGINE_STAGED = 1, FDTD_ENGINE_FUSED = 2

    type, bind(c) :: fdtd_grid
        integer(c_int32_t) :: nx, ny, nz, precision
        integer(c_int32_t) :: periodic(3), engine
        real(c_double) :: ds, courant
    end type fdtd_grid

    type, bind(c) :: fdtd_pml
        integer(c_int32_t) :: mode, npml(3)
        real(c_double) :: order, R0, sigmam, eps_low, eps_high
    end type fdtd_pml

    type, bind(c) :: fdtd_region
        integer(c_int32_t) :: kind, axis, lo(3), hi(3)
        real(c_double) :: center(3), radius, half_length
    end type fdtd_region

    type, bind(c) :: fdtd_drude2cp
        real(c_double) :: eps_inf, gam, omega_d
        real(c_double) :: cp_a(2), cp_phi(2), cp_omega(2), cp_gamma(2)
    end type fdtd_drude2cp

    type, bind(c) :: fdtd_graphene
        integer(c_int32_t) :: model, normal_axis, ref_init_order, reserved
        real(c_double) :: thickness, eps_inf, gam, omega_d
        real(c_double) :: cp_a(4), cp_phi(4), cp_omega(4), cp_gamma(4)
        real(c_double) :: fermi_ev, tau, temperature, wNL, dnuNL, xi3
    end type fdtd_graphene

    type, bind(c) :: fdtd_pulse
        real(c_double) :: t0, st, w0, amplitude
    end type fdtd_pulse

    ! the Lorentz-Jp family (mod3/goldLorentzJp_mod.f90 and siblings)
    type, bind(c) :: fdtd_lorentz_jp
        integer(c_int32_t) :: npols, jp_mask, nonlinear, ez_uses_jx
        real(c_double) :: eps_inf
        real(c_double) :: xi(4), w0(4), dnu(4)
        real(c_double) :: wNL, dnuNL, xi3
    end type fdtd_lorentz_jp
    ! TFSF plane wave (sim1/program.f90:86-91,217-262)
    type, bind(c) :: fdtd_tfsf
        integer(c_int32_t) :: axis, comp, lo, hi, src_index, npml
        real(c_double) :: eps, order, R0
        type(fdtd_pulse) :: pulse
    end type fdtd_tfsf
    integer(c_int32_t), parameter :: FDTD_LJP_GOLD = 0, FDTD_LJP_SI = 1, FDTD_LJP_TITANIUM = 2
    integer(c_int32_t), parameter :: FDTD_LJP_FILM_GOLD = 3, FDTD_LJP_GRAPHENE = 4
    integer(c_int32_t), parameter :: FDTD_LJP_GRAPHENE3_1NM = 5, FDTD_LJP_GRAPHENE_ND = 6

    interface
        subroutine fdtd_drude2cp_gold(p) bind(c, name="fdtd_drude2cp_gold")
            import :: fdtd_drude2cp
            type(fdtd_drude2cp), intent(out) :: p
        end subroutine
        subroutine fdtd_graphene_ref(p) bind(c, name="fdtd_graphene_ref")
            import :: fdtd_graphene
            type(fdtd_graphene), intent(out) :: p
        end subroutine
        integer(c_int) function fdtd_create(sim, grid) bind(c, name="fdtd_create")
            import :: c_int, c_ptr, fdtd_grid
            type(c_ptr), intent(out) :: sim
            type(fdtd_grid), intent(in) :: grid
        end function
        integer(c_int) function fdtd_destroy(sim) bind(c, name="fdtd_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: sim
        end function
        type(c_ptr) function fdtd_last_error() bind(c, name="fdtd_last_error")
            import :: c_ptr
        end function
        integer(c_int) function fdtd_set_pml(sim, pml) bind(c, name="fdtd_set_pml")
            import :: c_int, c_ptr, fdtd_pml
            type(c_ptr), value :: sim
            type(fdtd_pml), intent(in) :: pml
        end function
        integer(c_int) function fdtd_add_dielectric(sim, region, eps_r) bind(c, name="fdtd_add_dielectric")
            import :: c_int, c_ptr, c_double, fdtd_region
            type(c_ptr), value :: sim
            type(fdtd_region), intent(in) :: region
            real(c_double), value :: eps_r
        end function
        integer(c_int) function fdtd_add_drude2cp(sim, region, p) bind(c, name="fdtd_add_drude2cp")
            import :: c_int, c_ptr, fdtd_region, fdtd_drude2cp
            type(c_ptr), value :: sim
            type(fdtd_region), intent(in) :: region
            type(fdtd_drude2cp), intent(in) :: p
        end function
        integer(c_int) function fdtd_add_graphene(sim, region, p) bind(c, name="fdtd_add_graphene")
            import :: c_int, c_ptr, fdtd_region, fdtd_graphene
            type(c_ptr), value :: sim
            type(fdtd_region), intent(in) :: region
            type(fdtd_graphene), intent(in) :: p
        end function
        integer(c_int) function fdtd_add_lorentz_jp(sim, region, p) bind(c, name="fdtd_add_lorentz_jp")
            import :: c_int, c_ptr, fdtd_region, fdtd_lorentz_jp
            type(c_ptr), value :: sim
            type(fdtd_region), intent(in) :: region
            type(fdtd_lorentz_jp), intent(in) :: p
        end function
        integer(c_int) function fdtd_add_lorentz_d(sim, region, p) bind(c, name="fdtd_add_lorentz_d")
            import :: c_int, c_ptr, fdtd_region, fdtd_lorentz_jp
            type(c_ptr), value :: sim
            type(fdtd_region), intent(in) :: region
            type(fdtd_lorentz_jp), intent(in) :: p
        end function
        integer(c_int) function fdtd_lorentz_jp_preset(kind, p) bind(c, name="fdtd_lorentz_jp_preset")
            import :: c_int, c_int32_t, fdtd_lorentz_jp
            integer(c_int32_t), value :: kind
            type(fdtd_lorentz_jp), intent(out) :: p
        end function
        integer(c_int) function fdtd_add_tfsf(sim, t, id) bind(c, name="fdtd_add_tfsf")
            import :: c_int, c_ptr, fdtd_tfsf
            type(c_ptr), value :: sim
            type(fdtd_tfsf), intent(in) :: t
            integer(c_int), intent(out) :: id
        end function
        integer(c_int) function fdtd_add_tfsf_probe(sim, tid, idx, id) bind(c, name="fdtd_add_tfsf_probe")
            import :: c_int, c_ptr
            type(c_ptr), value :: sim
            integer(c_int), value :: tid, idx
            integer(c_int), intent(out) :: id
        end function
        integer(c_int) function fdtd_get_tfsf_line(sim, tid, e, h) bind(c, name="fdtd_get_tfsf_line")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: sim
            integer(c_int), value :: tid
            real(c_double), intent(out) :: e(*), h(*)
        end function
        integer(c_int) function fdtd_add_plane_source(sim, axis, idx, comp, p) bind(c, name="fdtd_add_plane_source")
            import :: c_int, c_ptr, fdtd_pulse
            type(c_ptr), value :: sim
            integer(c_int), value :: axis, idx, comp
            type(fdtd_pulse), intent(in) :: p
        end function
        integer(c_int) function fdtd_add_dft_plane(sim, axis, idx, nfreq, omega, comp_mask, n_begin, n_end, id) &
                bind(c, name="fdtd_add_dft_plane")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: sim
            integer(c_int), value :: axis, idx, nfreq, comp_mask, n_begin, n_end
            real(c_double), intent(in) :: omega(*)
            integer(c_int), intent(out) :: id
        end function
        integer(c_int) function fdtd_add_probe(sim, i, j, k, id) bind(c, name="fdtd_add_probe")
            import :: c_int, c_ptr
            type(c_ptr), value :: sim
            integer(c_int), value :: i, j, k
            integer(c_int), intent(out) :: id
        end function
        integer(c_int) function fdtd_commit(sim) bind(c, name="fdtd_commit")
            import :: c_int, c_ptr
            type(c_ptr), value :: sim
        end function
        integer(c_int) function fdtd_step(sim, nsteps) bind(c, name="fdtd_step")
            import :: c_int, c_ptr
            type(c_ptr), value :: sim
            integer(c_int), value :: nsteps
        end function
        integer(c_int) function fdtd_sync(sim) bind(c, name="fdtd_sync")
            import :: c_int, c_ptr
            type(c_ptr), value :: sim
        end function
        integer(c_int) function fdtd_checkpoint_size(sim, nbytes) bind(c, name="fdtd_checkpoint_size")
            import :: c_int, c_ptr, c_size_t
            type(c_ptr), value :: sim
            integer(c_size_t), intent(out) :: nbytes
        end function
        integer(c_int) function fdtd_checkpoint_save(sim, buf, nbytes) bind(c, name="fdtd_checkpoint_save")
            import :: c_int, c_ptr, c_size_t
            type(c_ptr), value :: sim, buf
            integer(c_size_t), value :: nbytes
        end function
        integer(c_int) function fdtd_checkpoint_load(sim, buf, nbytes) bind(c, name="fdtd_checkpoint_load")
            import :: c_int, c_ptr, c_size_t
            type(c_ptr), value :: sim, buf
            integer(c_size_t), value :: nbytes
        end function
        integer(c_int) function fdtd_get_dt(sim, dt) bind(c, name="fdtd_get_dt")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: sim
            real(c_double), intent(out) :: dt
        end function
        integer(c_int) function fdtd_get_field(sim, comp, dst) bind(c, name="fdtd_get_field")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: sim
            integer(c_int), value :: comp
            real(c_double), intent(out) :: dst(*)
        end function
        integer(c_int) function fdtd_get_dft(sim, id, comp, re, im) bind(c, name="fdtd_get_dft")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: sim
            integer(c_int), value :: id, comp
            real(c_double), intent(out) :: re(*), im(*)
        end function
        integer(c_int) function fdtd_get_flux(sim, id, S) bind(c, name="fdtd_get_flux")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: sim
            integer(c_int), value :: id
            real(c_double), intent(out) :: S(*)
        end function
        integer(c_int) function fdtd_get_probe_count(sim, id, nrec) bind(c, name="fdtd_get_probe_count")
            import :: c_int, c_ptr
            type(c_ptr), value :: sim
            integer(c_int), value :: id
            integer(c_int), intent(out) :: nrec
        end function
        integer(c_int) function fdtd_get_probe(sim, id, out) bind(c, name="fdtd_get_probe")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: sim
            integer(c_int), value :: id
            real(c_double), intent(out) :: out(*)
        end function
        integer(c_int) function fdtd_energy(sim, we, wh) bind(c, name="fdtd_energy")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: sim
            real(c_double), intent(out) :: we, wh
        end function
    end interface

contains

    ! fdtd_last_error() as a Fortran string
    function fdtd_error_text() result(msg)
        character(len=256) :: msg
        character(kind=c_char), pointer :: p(:)
        integer :: n
        msg = ' '
        call c_f_pointer(fdtd_last_error(), p, [256])
        do n = 1, 256
            if (p(n) == c_null_char) exit
            msg(n:n) = p(n)
        end do
    end function fdtd_error_text

    subroutine fdtd_check(rc, what)
        integer(c_int), intent(in) :: rc
        character(*), intent(in) :: what
        if (rc /= FDTD_OK) then
            print *, 'fdtd error in ', what, ': ', trim(fdtd_error_text())
            error stop 2
        end if
    end subroutine fdtd_check

    function box_region(xinit, xend, yinit, yend, zinit, zend) result(r)
        integer, intent(in) :: xinit, xend, yinit, yend, zinit, zend
        type(fdtd_region) :: r
        r%kind = FDTD_REGION_BOX
        r%axis = 0
        r%lo = [xinit, yinit, zinit]
        r%hi = [xend, yend, zend]
        r%center = 0.d0
        r%radius = 0.d0
        r%half_length = 0.d0
    end function box_region
end module fdtd_c
