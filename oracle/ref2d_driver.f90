! ref2d_driver.f90 -- single-domain driver around the reference's own 2D modules.
!
! TEST INFRASTRUCTURE ONLY (writes golden vectors; never part of the product).
!
! It is compiled against the reference's module sources where they lie
! (modules/src/misc/*.f90, modules/src/2D/{material,simpledielectric,gold,
! graphene,siLorentzJp,filmGoldLorentzJp,graphene3LorentzJp1nm}_mod.f90) and
! textually includes the reference's 2D Yee kernels
! (FDTDsimulations/2D/air-nanoAntennas_Si02-Sisubstrate/include/myfdtd_2Dfunc.f90),
! as sim2/program.f90:535 does.  Everything else restates the driver logic
! of sim2/program.f90:320-438 for one rank, with no PML (PEC y walls) or
! with the reference's own y-slab PML module (mod2/BERpmly_HzExEy_mod.f90,
! one rank: the object carries both slabs, as ycord = 0 = Ymax-1), the main
! updates then running over rows npmly..N_y-npmly as the reference's local
! range (sim2/include/mympi_2Dfunc.f90:124-131):
!   - the halo exchange of sim2/include/mympi_2Dfunc.f90 with PBC_x becomes
!     the x ghost copies it performs on one rank (Ey after E, Hz after H);
!   - materials are applied in the order given (last writer wins); the
!     Lorentz-Jp modules (update_EJp, E form from Hz) are applied in the same
!     pass, so test cases keep their cells apart from the other materials;
!   - the Dx line source of :331-341; point traces (n*dt, Ex, Hz) after each
!     step (:376,397,417).
! Usage: ref2d <config.nml> <outdir>
program ref2d
    use constants_mod
    use mydatatypes
    use material_mod
    use simpleDielectric_mod
    use gold_mod
    use graphene_mod
    use siLorentzJp_mod
    use filmgoldLorentzJp_mod
    use graphene3LorentzJp1nm_mod
    use BERpmly_HzExEy_mod
    implicit none
    integer :: nx, ny, nsteps
    real*8 :: ds, Ku, dt
    integer, parameter :: MAXM = 12
    integer :: nmat, mtype(MAXM), mbox(4, MAXM)
    real*8 :: meps(MAXM), gwNL(MAXM), gdnuNL(MAXM), gxi3(MAXM)
    logical :: s_on
    integer :: s_index
    real*8 :: s_t0, s_st, s_w0
    integer :: nprobe, p_ij(2, 8)
    integer :: npmlx, npmly, j0, j1
    real*8 :: order, R0, eps_low, eps_high, sigmam
    type(pmly) :: mypmly
    namelist /grid/ nx, ny, ds, Ku, nsteps
    namelist /pml/ npmlx, npmly, order, R0, eps_low, eps_high
    namelist /mats/ nmat, mtype, mbox, meps, gwNL, gdnuNL, gxi3
    namelist /src/ s_on, s_index, s_t0, s_st, s_w0
    namelist /probes/ nprobe, p_ij

    real*8, allocatable, dimension(:,:), target :: Hz, Ex, Ey, Dx, Dy
    real*8, dimension(:,:), pointer :: PEx, PEy, PDx, PDy, PHz
    type(ConstantDielectric) :: diel(MAXM)
    type(Gold) :: au(MAXM)
    type(Graphene) :: gr(MAXM)
    ! Lorentz-Jp modules: mtype 5 si, 7 film gold, 9 graphene3 1 nm (the 3D
    ! driver's numbering, FDTD_LJP_* + 4)
    type(si_LJp) :: lsi(MAXM)
    type(filmgold_LJp) :: lfilm(MAXM)
    type(SLG_3LJp1nm) :: lgr3(MAXM)
    type(Irange) :: absIx, absIy
    real*8, allocatable :: prb(:,:,:)
    real*8 :: t, gauss_t
    integer :: n, i, j, q, p
    character(len=512) :: cfg, outdir

    call get_command_argument(1, cfg)
    call get_command_argument(2, outdir)
    nmat = 0; mtype = 0; mbox = 0; meps = 1.d0; gwNL = 0.d0; gdnuNL = 0.d0; gxi3 = 0.d0
    s_on = .false.; s_index = 1; s_t0 = 0.d0; s_st = 1.d0; s_w0 = 0.d0
    nprobe = 0; p_ij = 1
    npmlx = 0; npmly = 0; order = 3.d0; R0 = dexp(-16.d0); eps_low = 1.d0; eps_high = 1.d0
    open(unit=7, file=trim(cfg), status='old')
    read(7, nml=grid)
    read(7, nml=pml)
    read(7, nml=mats)
    read(7, nml=src)
    read(7, nml=probes)
    close(7)

    dt = Ku*ds/c0
    allocate(Hz(0:nx+1,0:ny+1), Ex(0:nx+1,0:ny+1), Ey(0:nx+1,0:ny+1), Dx(0:nx+1,0:ny+1), Dy(0:nx+1,0:ny+1))
    Hz = 0.d0; Ex = 0.d0; Ey = 0.d0; Dx = 0.d0; Dy = 0.d0
    PEx => Ex; PEy => Ey; PDx => Dx; PDy => Dy; PHz => Hz

    absIx%init = 1; absIx%ends = nx
    absIy%init = 1; absIy%ends = ny
    j0 = 1; j1 = ny
    if (npmly > 0) then
        ! sim2/program.f90:93,237 (one rank: coords 0, Ndim 1)
        sigmam = (order+1.d0)*c0*log(1.d0/R0)/(2.d0*npmly*ds)
        call mypmly%initialize(eps_low*ep0, eps_high*ep0, sigmam, (/ npmlx, npmly /), order, dt, ds, &
                               (/ 0, 0 /), (/ nx, ny /), (/ 1, 1 /))
        j0 = npmly; j1 = ny-npmly
    end if
    do q = 1, nmat
        select case (mtype(q))
        case (1)
            call setbox(diel(q)%material, mbox(:, q))
            diel(q)%constant_ep = meps(q)
            call diel(q)%initialize_material(absIx, absIy, (/ nx, ny /))
        case (2)
            call setbox(au(q)%material, mbox(:, q))
            au(q)%dt = dt
            call au(q)%initialize_material(absIx, absIy, (/ nx, ny /))
        case (3)
            call setbox(gr(q)%material, mbox(:, q))
            gr(q)%dt = dt
            gr(q)%ds = ds
            gr(q)%wNL = gwNL(q)
            gr(q)%dnuNL = gdnuNL(q)
            gr(q)%xi3 = gxi3(q)
            call gr(q)%initialize_material(absIx, absIy, (/ nx, ny /))
        case (5)
            call setbox(lsi(q)%material, mbox(:, q))
            lsi(q)%dt = dt; lsi(q)%ds = ds
            call lsi(q)%initialize_material(absIx, absIy, (/ nx, ny /))
        case (7)
            call setbox(lfilm(q)%material, mbox(:, q))
            lfilm(q)%dt = dt; lfilm(q)%ds = ds
            call lfilm(q)%initialize_material(absIx, absIy, (/ nx, ny /))
        case (9)
            call setbox(lgr3(q)%material, mbox(:, q))
            lgr3(q)%dt = dt; lgr3(q)%ds = ds
            lgr3(q)%wNL = gwNL(q); lgr3(q)%dnuNL = gdnuNL(q); lgr3(q)%xi3 = gxi3(q)
            call lgr3(q)%initialize_material(absIx, absIy, (/ nx, ny /))
        end select
    end do

    allocate(prb(3, nsteps, max(nprobe, 1)))
    prb = 0.d0
    do n = 1, nsteps
        t = real(n-1, 8)*dt
        call updateD(1, nx, j0, j1)
        if (s_on) then
            gauss_t = dexp(-0.5d0*((t-s_t0)/s_st)**2)*dsin(s_w0*t)
            do i = 1, nx
                Dx(i, s_index) = Dx(i, s_index)+gauss_t
            end do
        end if
        do q = 1, nmat
            select case (mtype(q))
            case (1)
                call diel(q)%update_E(PEx, PEy, PDx, PDy)
            case (2)
                call au(q)%update_E(PEx, PEy, PDx, PDy)
            case (3)
                call gr(q)%update_E(PEx, PEy, PDx, PDy)
            case (5)
                call lsi(q)%update_EJp(PEx, PEy, PHz)
            case (7)
                call lfilm(q)%update_EJp(PEx, PEy, PHz)
            case (9)
                call lgr3(q)%update_EJp(PEx, PEy, PHz)
            end select
        end do
        if (npmly > 0) call mypmly%updateE(PEx, PHz)
        Ey(nx+1, :) = Ey(1, :)
        Ey(0, :) = Ey(nx, :)
        call updateH(1, nx, j0, j1)
        if (npmly > 0) call mypmly%updateH(PHz)
        Hz(nx+1, :) = Hz(1, :)
        Hz(0, :) = Hz(nx, :)
        do p = 1, nprobe
            i = p_ij(1, p); j = p_ij(2, p)
            prb(1, n, p) = n*dt
            prb(2, n, p) = Ex(i, j)
            prb(3, n, p) = Hz(i, j)
        end do
    end do

    open(unit=20, file=trim(outdir)//'/fields.bin', access='stream', form='unformatted', status='replace')
    write(20) Ex(1:nx, 1:ny), Ey(1:nx, 1:ny), Hz(1:nx, 1:ny), Dx(1:nx, 1:ny), Dy(1:nx, 1:ny)
    close(20)
    open(unit=22, file=trim(outdir)//'/probes.bin', access='stream', form='unformatted', status='replace')
    write(22) prb
    close(22)
    print *, 'done', nsteps, dt

contains
    include 'myfdtd_2Dfunc.f90'

    subroutine setbox(mt, b)
        type(material), intent(inout) :: mt
        integer, intent(in) :: b(4)
        mt%xinit = b(1); mt%xend = b(2)
        mt%yinit = b(3); mt%yend = b(4)
    end subroutine setbox
end program ref2d
