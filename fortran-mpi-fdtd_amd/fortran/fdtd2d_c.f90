! fdtd2d_c.f90 -- iso_c_binding interface to include/fdtd2d.h for Fortran
! hosts.  The material, pulse and status definitions shared with the 3D ABI
! come from fdtd_c.f90.
module fdtd2d_c
    use iso_c_binding
    use fdtd_c
    implicit none

    integer(c_int), parameter :: FDTD2D_TEZ = 0, FDTD2D_TMZ = 1
    integer(c_int), parameter :: FDTD2D_EX = 0, FDTD2D_EY = 1, FDTD2D_HZ = 2, FDTD2D_DX = 3, FDTD2D_DY = 4
    integer(c_int), parameter :: FDTD2D_EZ = 5, FDTD2D_HX = 6, FDTD2D_HY = 7, FDTD2D_DZ = 8
    integer(c_int), parameter :: FDTD2D_REGION_BOX = 0, FDTD2D_REGION_DISK = 1
    integer(c_int), parameter :: FDTD2D_DFT_EX = 1, FDTD2D_DFT_HZ = 4, FDTD2D_DFT_EZ = 8, FDTD2D_DFT_HX = 16

    type, bind(c) :: fdtd2d_grid
        integer(c_int32_t) :: nx, ny, precision, periodic_x
        real(c_double) :: ds, courant
        integer(c_int32_t) :: polarization = FDTD2D_TEZ, reserved = 0
    end type fdtd2d_grid

    type, bind(c) :: fdtd2d_pml
        integer(c_int32_t) :: npml(2)
        real(c_double) :: order, R0, sigmam, eps_low, eps_high
    end type fdtd2d_pml

    type, bind(c) :: fdtd2d_tfsf
        integer(c_int32_t) :: lo, hi, src_index, npml
        real(c_double) :: eps, order, R0
        type(fdtd_pulse) :: pulse
    end type fdtd2d_tfsf

    type, bind(c) :: fdtd2d_region
        integer(c_int32_t) :: kind, reserved, lo(2), hi(2)
        real(c_double) :: center(2), radius
    end type fdtd2d_region

    interface
        integer(c_int) function fdtd2d_create(sim, grid) bind(c, name="fdtd2d_create")
            import :: c_int, c_ptr, fdtd2d_grid
            type(c_ptr), intent(out) :: sim
            type(fdtd2d_grid), intent(in) :: grid
        end function
        integer(c_int) function fdtd2d_destroy(sim) bind(c, name="fdtd2d_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: sim
        end function
        integer(c_int) function fdtd2d_set_pml(sim, pml) bind(c, name="fdtd2d_set_pml")
            import :: c_int, c_ptr, fdtd2d_pml
            type(c_ptr), value :: sim
            type(fdtd2d_pml), intent(in) :: pml
        end function
        integer(c_int) function fdtd2d_add_dielectric(sim, region, eps_r) bind(c, name="fdtd2d_add_dielectric")
            import :: c_int, c_ptr, c_double, fdtd2d_region
            type(c_ptr), value :: sim
            type(fdtd2d_region), intent(in) :: region
            real(c_double), value :: eps_r
        end function
        integer(c_int) function fdtd2d_add_drude2cp(sim, region, p) bind(c, name="fdtd2d_add_drude2cp")
            import :: c_int, c_ptr, fdtd2d_region, fdtd_drude2cp
            type(c_ptr), value :: sim
            type(fdtd2d_region), intent(in) :: region
            type(fdtd_drude2cp), intent(in) :: p
        end function
        integer(c_int) function fdtd2d_add_graphene(sim, region, p) bind(c, name="fdtd2d_add_graphene")
            import :: c_int, c_ptr, fdtd2d_region, fdtd_graphene
            type(c_ptr), value :: sim
            type(fdtd2d_region), intent(in) :: region
            type(fdtd_graphene), intent(in) :: p
        end function
        integer(c_int) function fdtd2d_add_line_source(sim, j, p) bind(c, name="fdtd2d_add_line_source")
            import :: c_int, c_ptr, fdtd_pulse
            type(c_ptr), value :: sim
            integer(c_int), value :: j
            type(fdtd_pulse), intent(in) :: p
        end function
        integer(c_int) function fdtd2d_add_dft_line(sim, j, nfreq, omega, comp_mask, n_begin, n_end, id) &
                bind(c, name="fdtd2d_add_dft_line")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: sim
            integer(c_int), value :: j, nfreq, comp_mask, n_begin, n_end
            real(c_double), intent(in) :: omega(*)
            integer(c_int), intent(out) :: id
        end function
        integer(c_int) function fdtd2d_add_probe(sim, i, j, id) bind(c, name="fdtd2d_add_probe")
            import :: c_int, c_ptr
            type(c_ptr), value :: sim
            integer(c_int), value :: i, j
            integer(c_int), intent(out) :: id
        end function
        integer(c_int) function fdtd2d_add_lorentz_jp(sim, region, p) bind(c, name="fdtd2d_add_lorentz_jp")
            import :: c_int, c_ptr, fdtd2d_region, fdtd_lorentz_jp
            type(c_ptr), value :: sim
            type(fdtd2d_region), intent(in) :: region
            type(fdtd_lorentz_jp), intent(in) :: p
        end function
        integer(c_int) function fdtd2d_add_tfsf(sim, t, id) bind(c, name="fdtd2d_add_tfsf")
            import :: c_int, c_ptr, fdtd2d_tfsf
            type(c_ptr), value :: sim
            type(fdtd2d_tfsf), intent(in) :: t
            integer(c_int), intent(out) :: id
        end function
        integer(c_int) function fdtd2d_get_tfsf_line(sim, tid, e, h) bind(c, name="fdtd2d_get_tfsf_line")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: sim
            integer(c_int), value :: tid
            real(c_double), intent(out) :: e(*), h(*)
        end function
        integer(c_int) function fdtd2d_commit(sim) bind(c, name="fdtd2d_commit")
            import :: c_int, c_ptr
            type(c_ptr), value :: sim
        end function
        integer(c_int) function fdtd2d_step(sim, nsteps) bind(c, name="fdtd2d_step")
            import :: c_int, c_ptr
            type(c_ptr), value :: sim
            integer(c_int), value :: nsteps
        end function
        integer(c_int) function fdtd2d_sync(sim) bind(c, name="fdtd2d_sync")
            import :: c_int, c_ptr
            type(c_ptr), value :: sim
        end function
        integer(c_int) function fdtd2d_get_dt(sim, dt) bind(c, name="fdtd2d_get_dt")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: sim
            real(c_double), intent(out) :: dt
        end function
        integer(c_int) function fdtd2d_get_field(sim, comp, dst) bind(c, name="fdtd2d_get_field")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: sim
            integer(c_int), value :: comp
            real(c_double), intent(out) :: dst(*)
        end function
        integer(c_int) function fdtd2d_get_dft(sim, id, comp, re, im) bind(c, name="fdtd2d_get_dft")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: sim
            integer(c_int), value :: id, comp
            real(c_double), intent(out) :: re(*), im(*)
        end function
        integer(c_int) function fdtd2d_get_flux(sim, id, S) bind(c, name="fdtd2d_get_flux")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: sim
            integer(c_int), value :: id
            real(c_double), intent(out) :: S(*)
        end function
        integer(c_int) function fdtd2d_get_probe_count(sim, id, nrec) bind(c, name="fdtd2d_get_probe_count")
            import :: c_int, c_ptr
            type(c_ptr), value :: sim
            integer(c_int), value :: id
            integer(c_int), intent(out) :: nrec
        end function
        integer(c_int) function fdtd2d_get_probe(sim, id, out) bind(c, name="fdtd2d_get_probe")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: sim
            integer(c_int), value :: id
            real(c_double), intent(out) :: out(*)
        end function
    end interface

contains

    function box2_region(xinit, xend, yinit, yend) result(r)
        integer, intent(in) :: xinit, xend, yinit, yend
        type(fdtd2d_region) :: r
        r%kind = FDTD2D_REGION_BOX
        r%reserved = 0
        r%lo = [xinit, yinit]
        r%hi = [xend, yend]
        r%center = 0.d0
        r%radius = 0.d0
    end function box2_region
end module fdtd2d_c
