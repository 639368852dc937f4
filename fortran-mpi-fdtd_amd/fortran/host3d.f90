! host3d.f90 -- Fortran90 host for the gold-antenna/SiO2/Si 3D case, driving
! the gfx950 library through the C ABI (fdtd_c.f90).
!
! It keeps the structure of sim3/program.f90: compile-time parameters
! (:44-156), PML initialisation (:205), materials set up and applied in the
! update order air, Au, SiO2, Si (:230-274, :411-414), the incident-time
! window (:280-287), the frequency grid (:353-357), time-gated incident /
! reflected / transmitted monitors (:441-540), the frames every dn steps
! (:543-604) and the flux files (:613-685).
! The MPI layout (Xmax=1, Ymax=6, Zmax=1) becomes one domain on one GPU: the
! PML keeps the multi-rank semantics (y-low slab in air, y-high slab in Si),
! and the frame files keep the per-rank split (rank r holds global rows
! r*n_y+1 .. (r+1)*n_y).
!
! Usage: host3d [Nt [outdir [frames [dn]]]]
!   outdir: run directory holding transforms/, timesLiney/, timesGnuplotmap/
!           and timesVTK/ (created by the caller), default '.'
!   frames: bit 0 line frames (the reference's printTimesLine, on by
!           default), bit 1 map frames (printTimesMap), bit 2 VTK frames
!           (printTimesVTK); default 1
!   dn:     frame interval, default aint(4*tfinStop/20) as :285
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
    integer, parameter :: Ymax = 6, n_x = nx, n_y = ny/Ymax, n_z = nz
    integer :: ptty

    type(c_ptr) :: sim
    type(fdtd_grid) :: grid
    type(fdtd_pml) :: pml
    type(fdtd_drude2cp) :: au
    type(fdtd_pulse) :: pulse
    integer :: air_y(2), au_x(2), au_y(2), sio2_y(2), si_y(2)
    integer :: tfinStop, Nt, m, nrec, q, id_in, id_r, id_t, pr_in, pr_r, pr_t, mask
    integer :: frames, dn, n, frame
    real*8 :: freq, finit, fend, df, lambda, dt, time4incident, distance4incident
    real*8 :: arg(freqsteps), Sy(freqsteps)
    real*8, allocatable :: rec(:), fld(:, :)
    character(len=256) :: a1, outdir
    integer(c_int) :: rc, idc

    outdir = '.'
    if (command_argument_count() >= 2) call get_command_argument(2, outdir)
    frames = 1
    if (command_argument_count() >= 3) then
        call get_command_argument(3, a1)
        read (a1, *) frames
    end if

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
    dn = int(aint(Nt/20.d0))
    if (command_argument_count() >= 1) then
        call get_command_argument(1, a1)
        read(a1, *) Nt
    end if
    if (command_argument_count() >= 4) then
        call get_command_argument(4, a1)
        read (a1, *) dn
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

    ! the time loop of sim3/program.f90:382-607, one call per frame interval
    allocate (fld(nx*ny*nz, 0:5))
    n = 0
    frame = 0
    do while (iand(frames, 7) /= 0 .and. dn > 0 .and. n+dn <= Nt)
        rc = fdtd_step(sim, dn)
        call fdtd_check(rc, 'step')
        n = n+dn
        frame = frame+1
        do q = 0, 5
            call fdtd_check(fdtd_get_field(sim, q, fld(:, q)), 'field')
        end do
        call write_frame(frame)
    end do
    if (Nt > n) then
        rc = fdtd_step(sim, Nt-n)
        call fdtd_check(rc, 'step')
    end if
    call fdtd_check(fdtd_sync(sim), 'sync')

    ! flux spectra, sim3/program.f90:613-685
    call write_flux(id_in, 'r0Syin.dat')
    call write_flux(id_r, 'r0Syr.dat')
    call write_tf_r(id_r, 'r0tf_r.dat')
    write (a1, '(I4)') (ptty-1)/n_y
    call write_flux(id_t, 'r'//trim(adjustl(a1))//'Syt.dat')

    ! probe traces in the reference's '(7E30.19)' format
    call fdtd_check(fdtd_get_probe_count(sim, pr_in, idc), 'probe count')
    nrec = idc
    allocate(rec(7*max(nrec, 1)))
    call fdtd_check(fdtd_get_probe(sim, pr_in, rec), 'probe')
    open(unit=2, file=trim(outdir)//'/transforms/r0infld.dat', status='replace')
    open(unit=3, file=trim(outdir)//'/transforms/r0rfld.dat', status='replace')
    do q = 1, nrec
        if (q < tfinStop) write(2, '(7E30.19)') rec(7*(q-1)+1:7*q)
        if (q > tfinStop) write(3, '(7E30.19)') rec(7*(q-1)+1:7*q)
    end do
    close(2)
    close(3)
    call fdtd_check(fdtd_get_probe(sim, pr_t, rec), 'probe')
    write (a1, '(I4)') (ptty-1)/n_y
    open(unit=4, file=trim(outdir)//'/transforms/r'//trim(adjustl(a1))//'tfld.dat', status='replace')
    do q = 1, nrec
        write(4, '(7E30.19)') rec(7*(q-1)+1:7*q)
    end do
    close(4)
    call fdtd_check(fdtd_destroy(sim), 'destroy')
    print *, 'done'

contains
    ! one frame of sim3/program.f90:543-604 for every rank of the y split
    ! (Xmax = Zmax = 1: every rank holds the line x = ic, z = kc and the
    ! plane z = kc; the VTK frames come from the first and last y ranks)
    subroutine write_frame(fr)
        integer, intent(in) :: fr
        integer :: r, i, j, k, jg, c, u
        character(len=8) :: rs, fs
        character(len=512) :: fname
        write (fs, '(I4)') fr
        do r = 0, Ymax-1
            write (rs, '(I4)') r
            if (btest(frames, 2) .and. (r == 0 .or. r == Ymax-1)) then
                u = 100+fr
                fname = trim(outdir)//'/timesVTK/r'//trim(adjustl(rs))//'Ex_time'//trim(adjustl(fs))//'.vtk'
                open (unit=u, file=fname, status='unknown')
                call make_vtk_header(u, 'Ex', n_x, n_y, n_z)
                do k = 1, n_z
                    do j = 1, n_y
                        do i = 1, n_x
                            c = i+nx*(r*n_y+j-1)+nx*ny*(k-1)
                            write (u, '(E30.19)', advance='no') fld(c, FDTD_HZ)
                        end do
                    end do
                    write (u, *)
                end do
                close (u)
            end if
            if (btest(frames, 0)) then
                u = 400+fr
                fname = trim(outdir)//'/timesLiney/r'//trim(adjustl(rs))//'E_time'//trim(adjustl(fs))//'.dat'
                open (unit=u, file=fname, status='unknown')
                do j = 1, n_y
                    jg = r*n_y+j
                    c = ic+nx*(jg-1)+nx*ny*(kc-1)
                    write (u, '(7F10.5)') jg*ds/nano, fld(c, 0:5)
                end do
                close (u)
            end if
            if (btest(frames, 1)) then
                u = 500+fr
                fname = trim(outdir)//'/timesGnuplotmap/r'//trim(adjustl(rs))//'Ekc_time'//trim(adjustl(fs))//'.dat'
                open (unit=u, file=fname, status='unknown')
                do j = 1, n_y
                    jg = r*n_y+j
                    do i = 1, n_x
                        c = i+nx*(jg-1)+nx*ny*(kc-1)
                        write (u, '(8E30.19)') i*ds/nano, jg*ds/nano, fld(c, 0:5)
                    end do
                    write (u, *)
                end do
                close (u)
            end if
        end do
    end subroutine write_frame

    ! legacy VTK header of misc/constants_mod.f90:45-61
    subroutine make_vtk_header(u, msg, mx, my, mz)
        integer, intent(in) :: u, mx, my, mz
        character(*), intent(in) :: msg
        write (u, '(a)') '# vtk DataFile Version 2.0'
        write (u, '(a)') msg
        write (u, '(a)') 'ASCII'
        write (u, '(a)') 'DATASET STRUCTURED_POINTS'
        write (u, '(a,5I5)') 'DIMENSIONS', mx, my, mz
        write (u, '(a)') 'ASPECT_RATIO 1 1 1'
        write (u, '(a)') 'ORIGIN 0 0 0'
        write (u, '(a,I10)') 'POINT_DATA', mx*my*mz
        write (u, '(a)') 'SCALARS volume_Ez double 1'
        write (u, '(a)') 'LOOKUP_TABLE default'
    end subroutine make_vtk_header

    ! per-x flux products on the line z = kc of the reflected plane,
    ! sim3/program.f90:637-661 (one block of n_x lines per frequency)
    subroutine write_tf_r(id, fname)
        integer, intent(in) :: id
        character(*), intent(in) :: fname
        real*8, allocatable :: re(:, :), im(:, :)
        integer :: mm, i, c, cc(4)
        cc = [FDTD_EX, FDTD_EZ, FDTD_HX, FDTD_HZ]
        allocate (re(nx*nz*freqsteps, 4), im(nx*nz*freqsteps, 4))
        do i = 1, 4
            call fdtd_check(fdtd_get_dft(sim, id, cc(i), re(:, i), im(:, i)), 'dft')
        end do
        open (unit=21, file=trim(outdir)//'/transforms/'//fname, status='replace')
        do mm = 1, freqsteps
            lambda = 2.d0*pi*c0/arg(mm)
            do i = 1, nx
                c = i+nx*(kc-1)+nx*nz*(mm-1)
                write (21, '(2F15.5,4E20.10)') lambda/nano, i*ds/nano, re(c, 3)*re(c, 2), im(c, 3)*im(c, 2), &
                    re(c, 4)*re(c, 1), im(c, 4)*im(c, 1)
            end do
            write (21, *)
        end do
        close (21)
    end subroutine write_tf_r

    subroutine write_flux(id, fname)
        integer, intent(in) :: id
        character(*), intent(in) :: fname
        integer :: mm
        call fdtd_check(fdtd_get_flux(sim, id, Sy), 'flux')
        open(unit=10, file=trim(outdir)//'/transforms/'//fname, status='replace')
        do mm = 1, freqsteps
            lambda = 2.d0*pi*c0/arg(mm)
            write(10, *) lambda/nano, Sy(mm)
        end do
        close(10)
    end subroutine write_flux
end program host3d
