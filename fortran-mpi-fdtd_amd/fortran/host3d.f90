! host3d.f90 -- Fortran90 host for the gold-antenna/SiO2/Si 3D case, driving
! the gfx950 library through the C ABI (fdtd_c.f90).
!
! It keeps the structure of sim3/program.f90: compile-time parameters
! (:44-156), PML initialisation (:205), materials set up and applied in the
! update order air, Au, SiO2, Si (:230-274, :411-414), the incident-time
! window (:280-287), the frequency grid (:353-357), time-gated incident /
! reflected / transmitted monitors (:441-540) and the flux files (:613-685).
! The MPI layout (Xmax=1, Ymax=6, Zmax=1) becomes one domain on one GPU: the
! PML keeps the multi-rank semantics (y-low slab in air, y-high slab in Si).
!
! Usage: host3d [Nt_override [outdir]]
program host3d
    use iso_c_binding
    use fdtd_c
    implicit none
    real*8, parameter :: pi = 4.d0*datan(1.d0)
    real*8, parameter :: ep0 = 8.85418781762039d-12, mu0 = pi*4d-7
    real*8, parameter :: c0 = 1.d0/dsqrt(ep0*mu0), nano = 1.d-9
    real*8, parameter :: lambda0 = 633.d0*nano, f0 = c0/lambda0
    real*8, parameter :: Ku = 0.5d0
    real*8, parameter :: ds = 3.999d0*nano
    real*8, parameter :: xreal = 500.d0*nano, yreal = 5064.d0*nano, zreal = 25.d0*ds
    integer, parameter :: nx = floor(xreal/ds), ny = floor(yreal/ds), nz = floor(zreal/ds)
    integer, parameter :: ic = nx/2, kc = nz/2
    integer, parameter :: npmlx = 10, npmly = 10, npmlz = 10
    real*8, parameter :: order = 3.d0, R0 = dexp(-16.d0)
    real*8, parameter :: n_SiO2 = 1.46d0, n_Si = 3.88d0, n_air = 1.d0
    real*8, parameter :: periodoin = 1.d0/f0, w0 = 2.d0*pi*f0, beta = 1.d0
    real*8, parameter :: st = beta*periodoin, t0 = 4.d0*st
    integer, parameter :: sposy = npmly+2
    integer, parameter :: freqsteps = 100
    real*8, parameter :: lambdainit = 500d-9, lambdaend = 800d-9
    integer, parameter :: ptinx = ic, ptiny = sposy+2, ptinz = kc
    integer :: ptty

    type(c_ptr) :: sim
    type(fdtd_grid) :: grid
    type(fdtd_pml) :: pml
    type(fdtd_drude2cp) :: au
    type(fdtd_pulse) :: pulse
    integer :: air_y(2), au_x(2), au_y(2), sio2_y(2), si_y(2)
    integer :: tfinStop, Nt, m, nrec, q, id_in, id_r, id_t, pr_in, pr_r, pr_t, mask
    real*8 :: freq, finit, fend, df, lambda, dt, time4incident, distance4incident
    real*8 :: arg(freqsteps), Sy(freqsteps)
    real*8, allocatable :: rec(:)
    character(len=256) :: a1, outdir
    integer(c_int) :: rc, idc

    outdir = 'transforms'
    if (command_argument_count() >= 2) call get_command_argument(2, outdir)

    grid%nx = nx; grid%ny = ny; grid%nz = nz
    grid%precision = FDTD_FP64
    grid%periodic = [1, 0, 1]           ! PBC_x, PBC_y, PBC_z (sim3/program.f90:12)
    grid%engine = FDTD_ENGINE_AUTO
    grid%ds = ds
    grid%courant = Ku
    call fdtd_check(fdtd_create(sim, grid), 'create')
    call fdtd_check(fdtd_get_dt(sim, dt), 'get_dt')

    pml%mode = FDTD_PML_REF_Y
    pml%npml = [npmlx, npmly, npmlz]
    pml%order = order
    pml%R0 = R0
    pml%sigmam = 0.d0                   ! derived from R0 as sim3/program.f90:102
    pml%eps_low = n_air**2
    pml%eps_high = n_Si**2
    call fdtd_check(fdtd_set_pml(sim, pml), 'set_pml')

    ! material boxes, sim3/program.f90:230-275 (same integer truncations)
    air_y = [npmly+1, int(0.8*ny)]
    au_x = [int(ic-0.5*floor(100.d0*nano/ds)+1), int(ic+0.5*floor(100.d0*nano/ds))]
    au_y = [air_y(2)-floor(20.d0*nano/ds)+1, air_y(2)]
    sio2_y = [air_y(2)+1, air_y(2)+1+floor(284.d0*nano/ds)-1]
    si_y = [sio2_y(2)+1, ny-npmly]
    ptty = int(0.5*(si_y(2)+si_y(1)))
    call fdtd_drude2cp_gold(au)
    ! update order of sim3/program.f90:411-414: air, Au, SiO2, Si
    call fdtd_check(fdtd_add_dielectric(sim, box_region(1, nx, air_y(1), air_y(2), 1, nz), n_air**2), 'air')
    call fdtd_check(fdtd_add_drude2cp(sim, box_region(au_x(1), au_x(2), au_y(1), au_y(2), 1, nz), au), 'Au')
    call fdtd_check(fdtd_add_dielectric(sim, box_region(1, nx, sio2_y(1), sio2_y(2), 1, nz), n_SiO2**2), 'SiO2')
    call fdtd_check(fdtd_add_dielectric(sim, box_region(1, nx, si_y(1), si_y(2), 1, nz), n_Si**2), 'Si')

    ! time window, sim3/program.f90:280-284
    distance4incident = ds*real(air_y(2)-sposy, 8)+beta*4.5d0*lambda0/n_air
    time4incident = distance4incident/(c0/n_air)
    tfinStop = floor(time4incident/dt)
    Nt = 4*tfinStop
    if (command_argument_count() >= 1) then
        call get_command_argument(1, a1)
        read(a1, *) Nt
    end if

    pulse%t0 = t0; pulse%st = st; pulse%w0 = w0; pulse%amplitude = n_air
    call fdtd_check(fdtd_add_plane_source(sim, 1, sposy, FDTD_EX, pulse), 'source')

    finit = c0/lambdaend
    fend = c0/lambdainit
    df = (fend-finit)/freqsteps
    freq = finit
    do m = 1, freqsteps
        arg(m) = 2.d0*pi*freq
        freq = freq+df
    end do
    mask = ibset(ibset(ibset(ibset(0, FDTD_EX), FDTD_EZ), FDTD_HX), FDTD_HZ)
    call fdtd_check(fdtd_add_dft_plane(sim, 1, ptiny, freqsteps, arg, mask, 1, tfinStop-1, idc), 'dft in')
    id_in = idc
    call fdtd_check(fdtd_add_dft_plane(sim, 1, ptiny, freqsteps, arg, mask, tfinStop+1, 0, idc), 'dft r')
    id_r = idc
    call fdtd_check(fdtd_add_dft_plane(sim, 1, ptty, freqsteps, arg, mask, 1, 0, idc), 'dft t')
    id_t = idc
    call fdtd_check(fdtd_add_probe(sim, ptinx, ptiny, ptinz, idc), 'probe in')
    pr_in = idc
    pr_r = pr_in
    call fdtd_check(fdtd_add_probe(sim, ic, ptty, kc, idc), 'probe t')
    pr_t = idc
    call fdtd_check(fdtd_commit(sim), 'commit')
    print *, 'Nx=', nx, 'Ny=', ny, 'Nz=', nz, 'dt=', dt, 'Nt=', Nt, 'tfinStop=', tfinStop
    print *, 'Au:xi', au_x(1), 'Au:xf', au_x(2), 'Au:yi', au_y(1), 'Au:yf', au_y(2)

    ! the time loop of sim3/program.f90:382-607 is one call
    rc = fdtd_step(sim, Nt)
    call fdtd_check(rc, 'step')
    call fdtd_check(fdtd_sync(sim), 'sync')

    ! flux spectra, sim3/program.f90:613-685
    call write_flux(id_in, 'r0Syin.dat')
    call write_flux(id_r, 'r0Syr.dat')
    call write_flux(id_t, 'r0Syt.dat')

    ! probe traces in the reference's '(7E30.19)' format
    call fdtd_check(fdtd_get_probe_count(sim, pr_in, idc), 'probe count')
    nrec = idc
    allocate(rec(7*max(nrec, 1)))
    call fdtd_check(fdtd_get_probe(sim, pr_in, rec), 'probe')
    open(unit=2, file=trim(outdir)//'/r0infld.dat', status='replace')
    open(unit=3, file=trim(outdir)//'/r0rfld.dat', status='replace')
    do q = 1, nrec
        if (q < tfinStop) write(2, '(7E30.19)') rec(7*(q-1)+1:7*q)
        if (q > tfinStop) write(3, '(7E30.19)') rec(7*(q-1)+1:7*q)
    end do
    close(2)
    close(3)
    call fdtd_check(fdtd_get_probe(sim, pr_t, rec), 'probe')
    open(unit=4, file=trim(outdir)//'/r0tfld.dat', status='replace')
    do q = 1, nrec
        write(4, '(7E30.19)') rec(7*(q-1)+1:7*q)
    end do
    close(4)
    call fdtd_check(fdtd_destroy(sim), 'destroy')
    print *, 'done'

contains
    subroutine write_flux(id, fname)
        integer, intent(in) :: id
        character(*), intent(in) :: fname
        integer :: mm
        call fdtd_check(fdtd_get_flux(sim, id, Sy), 'flux')
        open(unit=10, file=trim(outdir)//'/'//fname, status='replace')
        do mm = 1, freqsteps
            lambda = 2.d0*pi*c0/arg(mm)
            write(10, *) lambda/nano, Sy(mm)
        end do
        close(10)
    end subroutine write_flux
end program host3d
