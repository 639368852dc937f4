! host2d.f90 -- Fortran90 host for the gold-strip/SiO2/Si 2D case, driving
! the gfx950 library through the C ABI (fdtd2d_c.f90).
!
! It keeps the structure of sim2/program.f90: parameters (:40-141),
! PML initialisation (:237), materials set up and applied in the update
! order air, Au, SiO2, Si (:171-214, :345-348), the incident-time window
! (:219-230), the frequency grid (:294-298), time-gated monitors and traces
! (:373-432), line and map frames every dn steps (:436-473), and the flux
! files (:479-529), and it writes the same files under the same names as
! the reference's 3-rank run (Xmax=1, Ymax=3): transforms/r0infld.dat,
! r0rfld.dat, r2tfld.dat, r0Syin.dat, r0Syr.dat, r0tf_r.dat, r2Syt.dat,
! timesLiney/r<rank>E_time<frame>.dat, timesGnuplotmap/r<rank>E_time<frame>.dat,
! and transforms/RTA_2Dfdtd.data as sim2/transforms/calR_Sy.f90 computes it
! (R = -Syr/Syin, T = Syt/Syin, A = 1 - R - T).  The MPI layout becomes one
! domain; rank numbers survive only in the file names and frame splits.
!
! Usage: host2d [Nt_override [outdir [maps]]]   (outdir must hold transforms/,
! timesLiney/ and timesGnuplotmap/; maps = 0 skips the map frames)
program host2d
    use iso_c_binding
    use fdtd_c
    use fdtd2d_c
    implicit none
    real*8, parameter :: pi = 4.d0*datan(1.d0)
    real*8, parameter :: ep0 = 8.85418781762039d-12, mu0 = pi*4d-7
    real*8, parameter :: c0 = 1.d0/dsqrt(ep0*mu0), nano = 1.d-9
    real*8, parameter :: lambda0 = 633.d0*nano, f0 = c0/lambda0
    real*8, parameter :: Ku = 0.5d0
    real*8, parameter :: ds = 3.999d0*nano
    real*8, parameter :: xreal = 500.d0*nano, yreal = 5064.d0*nano
    integer, parameter :: Xmax = 1, Ymax = 3
    integer, parameter :: nx = floor(xreal/ds), ny = floor(yreal/ds)
    integer, parameter :: ic = nx/2
    integer, parameter :: n_x = nx/Xmax, n_y = ny/Ymax
    integer, parameter :: npmlx = 10, npmly = 10
    real*8, parameter :: order = 3.d0, R0 = dexp(-16.d0)
    real*8, parameter :: sigmam = (order+1.d0)*c0*log(1.d0/R0)/(2.d0*npmly*ds)
    real*8, parameter :: n_SiO2 = 1.46d0, n_Si = 3.88d0, n_air = 1.d0
    real*8, parameter :: periodoin = 1.d0/f0, w0 = 2.d0*pi*f0, beta = 1.d0
    real*8, parameter :: st = beta*periodoin, t0 = 4.d0*st
    integer, parameter :: sposy = npmly+2
    integer, parameter :: freqsteps = 200
    real*8, parameter :: lambdainit = 500d-9, lambdaend = 800d-9
    real*8, parameter :: finit = c0/lambdaend, fend = c0/lambdainit
    real*8, parameter :: df = (fend-finit)/freqsteps
    integer, parameter :: ptinx = ic, ptiny = sposy+2, ptry = ptiny, pttx = ic

    type(c_ptr) :: sim
    type(fdtd2d_grid) :: grid
    type(fdtd2d_pml) :: pml
    type(fdtd_drude2cp) :: au
    type(fdtd_pulse) :: pulse
    integer :: air_y(2), au_x(2), au_y(2), sio2_y(2), si_y(2), ptty
    integer :: tfinStop, Nt, dn, m, i, j, r, frame, n, id_in, id_r, id_t, pr_in, pr_t, rk_t, maps
    real*8 :: freq, lambda, dt, time4incident, distance4incident, goldLength, goldThick
    real*8 :: arg(freqsteps), Syin(freqsteps), Syr(freqsteps), Syt(freqsteps)
    real*8, allocatable :: rec_in(:), rec_t(:), ex(:), hz(:), rex(:), iex(:), rhz(:), ihz(:)
    character(len=600) :: a1, outdir, fname
    character(len=8) :: rankstr, framestr
    integer(c_int) :: idc

    outdir = '.'
    maps = 1
    if (command_argument_count() >= 2) call get_command_argument(2, outdir)
    if (command_argument_count() >= 3) then
        call get_command_argument(3, a1)
        read (a1, *) maps
    end if

    grid%nx = nx; grid%ny = ny
    grid%precision = FDTD_FP64
    grid%periodic_x = 1                   ! PBC_x (sim2/program.f90:12)
    grid%ds = ds
    grid%courant = Ku
    grid%polarization = FDTD2D_TEZ        ! the reference's 2D driver is TEz
    call fdtd_check(fdtd2d_create(sim, grid), 'create')
    call fdtd_check(fdtd2d_get_dt(sim, dt), 'get_dt')

    ! materials, sim2/program.f90:171-214 (the same float-to-integer truncations)
    air_y = [npmly, int(0.8*ny)]
    goldLength = 100.d0*nano
    goldThick = 20.d0*nano
    au_x = [int(ic-0.5*floor(goldLength/ds)+1), int(ic+0.5*floor(goldLength/ds))]
    au_y = [air_y(2)-floor(goldThick/ds)+1, air_y(2)]
    sio2_y(1) = air_y(2)+1
    sio2_y(2) = sio2_y(1)+floor(284.d0*nano/ds)-1
    si_y = [sio2_y(2)+1, ny-npmly]
    ptty = int(0.5*(si_y(1)+si_y(2)))

    ! time window, :219-230
    distance4incident = ds*real(air_y(2)-sposy, 8)+beta*4.5d0*lambda0/n_air
    time4incident = distance4incident/(c0/n_air)
    tfinStop = floor(time4incident/dt)
    Nt = 4*tfinStop
    if (command_argument_count() >= 1) then
        call get_command_argument(1, a1)
        read (a1, *) Nt
    end if
    dn = int(aint(4*tfinStop/20.d0))

    pml%npml = [npmlx, npmly]
    pml%order = order
    pml%R0 = R0
    pml%sigmam = sigmam
    pml%eps_low = n_air**2
    pml%eps_high = n_Si**2
    call fdtd_check(fdtd2d_set_pml(sim, pml), 'pml')
    call fdtd_drude2cp_gold(au)
    call fdtd_check(fdtd2d_add_dielectric(sim, box2_region(1, nx, air_y(1), air_y(2)), n_air**2), 'air')
    call fdtd_check(fdtd2d_add_drude2cp(sim, box2_region(au_x(1), au_x(2), au_y(1), au_y(2)), au), 'Au')
    call fdtd_check(fdtd2d_add_dielectric(sim, box2_region(1, nx, sio2_y(1), sio2_y(2)), n_SiO2**2), 'SiO2')
    call fdtd_check(fdtd2d_add_dielectric(sim, box2_region(1, nx, si_y(1), si_y(2)), n_Si**2), 'Si')
    pulse%t0 = t0; pulse%st = st; pulse%w0 = w0; pulse%amplitude = 1.d0
    call fdtd_check(fdtd2d_add_line_source(sim, sposy, pulse), 'source')

    ! frequency grid by repeated addition, :294-298
    freq = finit
    do m = 1, freqsteps
        arg(m) = 2.d0*pi*freq
        freq = freq+df
    end do
    call fdtd_check(fdtd2d_add_dft_line(sim, ptiny, freqsteps, arg, FDTD2D_DFT_EX+FDTD2D_DFT_HZ, 1, tfinStop-1, idc), 'in')
    id_in = idc
    call fdtd_check(fdtd2d_add_dft_line(sim, ptry, freqsteps, arg, FDTD2D_DFT_EX+FDTD2D_DFT_HZ, tfinStop+1, 0, idc), 'r')
    id_r = idc
    call fdtd_check(fdtd2d_add_dft_line(sim, ptty, freqsteps, arg, FDTD2D_DFT_EX+FDTD2D_DFT_HZ, 1, 0, idc), 't')
    id_t = idc
    call fdtd_check(fdtd2d_add_probe(sim, ptinx, ptiny, idc), 'probe in')
    pr_in = idc
    call fdtd_check(fdtd2d_add_probe(sim, pttx, ptty, idc), 'probe t')
    pr_t = idc
    call fdtd_check(fdtd2d_commit(sim), 'commit')
    print *, 'Nx=', nx, 'Ny=', ny, 'dt=', dt, 'tfinStop=', tfinStop, 'Nt=', Nt, 'dn=', dn

    ! time loop in frames of dn steps, :320-473
    allocate (ex(nx*ny), hz(nx*ny))
    n = 0
    frame = 0
    do while (n+dn <= Nt .and. dn > 0)
        call fdtd_check(fdtd2d_step(sim, dn), 'step')
        n = n+dn
        frame = frame+1
        call fdtd_check(fdtd2d_get_field(sim, FDTD2D_EX, ex), 'Ex')
        call fdtd_check(fdtd2d_get_field(sim, FDTD2D_HZ, hz), 'Hz')
        write (framestr, '(I4)') frame
        do r = 0, Ymax-1
            write (rankstr, '(I4)') r
            fname = trim(outdir)//'/timesLiney/r'//trim(adjustl(rankstr))//'E_time'//trim(adjustl(framestr))//'.dat'
            open (unit=40, file=fname, status='unknown')
            do j = r*n_y+1, (r+1)*n_y
                write (40, '(3ES26.17E3)') j*ds/nano, ex(ic+nx*(j-1)), hz(ic+nx*(j-1))
            end do
            close (40)
            if (maps /= 0) then
                fname = trim(outdir)//'/timesGnuplotmap/r'//trim(adjustl(rankstr))//'E_time'// &
                        trim(adjustl(framestr))//'.dat'
                open (unit=50, file=fname, status='unknown')
                do j = r*n_y+1, (r+1)*n_y
                    do i = 1, n_x
                        write (50, '(3ES26.17E3)') i*1.d0, j*1.d0, hz(i+nx*(j-1))
                    end do
                    write (50, *)
                end do
                close (50)
            end if
        end do
    end do
    if (Nt > n) call fdtd_check(fdtd2d_step(sim, Nt-n), 'step')

    ! traces (:373-432): rank 0 holds ptiny, rank ptty/n_y holds ptty
    allocate (rec_in(3*Nt), rec_t(3*Nt))
    call fdtd_check(fdtd2d_get_probe(sim, pr_in, rec_in), 'probe in')
    call fdtd_check(fdtd2d_get_probe(sim, pr_t, rec_t), 'probe t')
    rk_t = (ptty-1)/n_y
    write (rankstr, '(I4)') rk_t
    open (unit=2, file=trim(outdir)//'/transforms/r0infld.dat', status='unknown')
    open (unit=3, file=trim(outdir)//'/transforms/r0rfld.dat', status='unknown')
    open (unit=4, file=trim(outdir)//'/transforms/r'//trim(adjustl(rankstr))//'tfld.dat', status='unknown')
    do n = 1, Nt
        if (n < tfinStop) write (2, '(3ES26.17E3)') rec_in(3*n-2), rec_in(3*n-1), rec_in(3*n)
        if (n > tfinStop) write (3, '(E30.19,4F30.19)') rec_in(3*n-2), rec_in(3*n-1), rec_in(3*n)
        write (4, '(E30.19,4F30.19)') rec_t(3*n-2), rec_t(3*n-1), rec_t(3*n)
    end do
    close (2); close (3); close (4)

    ! spectra (:479-529)
    call fdtd_check(fdtd2d_get_flux(sim, id_in, Syin), 'flux in')
    call fdtd_check(fdtd2d_get_flux(sim, id_r, Syr), 'flux r')
    call fdtd_check(fdtd2d_get_flux(sim, id_t, Syt), 'flux t')
    allocate (rex(nx*freqsteps), iex(nx*freqsteps), rhz(nx*freqsteps), ihz(nx*freqsteps))
    call fdtd_check(fdtd2d_get_dft(sim, id_r, FDTD2D_EX, rex, iex), 'dft r Ex')
    call fdtd_check(fdtd2d_get_dft(sim, id_r, FDTD2D_HZ, rhz, ihz), 'dft r Hz')
    open (unit=10, file=trim(outdir)//'/transforms/r0Syin.dat', status='unknown')
    open (unit=11, file=trim(outdir)//'/transforms/r0Syr.dat', status='unknown')
    open (unit=21, file=trim(outdir)//'/transforms/r0tf_r.dat', status='unknown')
    open (unit=12, file=trim(outdir)//'/transforms/r'//trim(adjustl(rankstr))//'Syt.dat', status='unknown')
    open (unit=5, file=trim(outdir)//'/transforms/RTA_2Dfdtd.data', status='unknown')
    do m = 1, freqsteps
        lambda = 2.d0*pi*c0/arg(m)
        write (10, *) lambda/nano, Syin(m)
        do i = 1, nx
            write (21, '(2F15.5,2E20.10)') lambda/nano, i*ds/nano, &
                rex(i+nx*(m-1))*rhz(i+nx*(m-1)), ihz(i+nx*(m-1))*iex(i+nx*(m-1))
        end do
        write (11, *) lambda/nano, Syr(m)
        write (21, *)
        write (12, *) lambda/nano, Syt(m)
        write (5, '(4ES26.17E3)') lambda/nano, -Syr(m)/Syin(m), Syt(m)/Syin(m), 1.d0-(-Syr(m)/Syin(m))-Syt(m)/Syin(m)
    end do
    close (10); close (11); close (12); close (21); close (5)
    call fdtd_check(fdtd2d_destroy(sim), 'destroy')
end program host2d
