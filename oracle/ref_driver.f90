! ref_driver.f90 -- single-domain driver around the reference's own 3D modules.
!
! TEST INFRASTRUCTURE ONLY (writes golden vectors; never part of the product).
!
! It is compiled against the reference's module sources where they lie
! (modules/src/misc/*.f90, modules/src/3D/{material,simpledielectric,gold,
! graphene,BERpmly}_mod.f90) and textually includes the reference's Yee
! kernels (FDTDsimulations/3D/air-AuAntennasSiO2-Si/include/myfdtd_func.f90),
! exactly as sim3/program.f90:689-691 does.  Everything else here restates the
! driver logic of sim3/program.f90 for one rank:
!   - the MPI Cartesian layout becomes one domain; the halo exchange of
!     sim3/include/mympi_func.f90:241-690 becomes the periodic ghost copies
!     it performs when Xmax = Zmax = 1 (send-to-self), and nothing along y;
!   - the y-PML is two pmly objects, one with ycord = 0 and one with
!     ycord = Ymax-1 (Ymax = 2): the multi-rank semantics of the reference;
!   - materials are applied in the order given (last writer wins);
!   - the Lorentz-Jp family (mod3/*LorentzJp*_mod.f90, titaniumJp_mod.f90,
!     NodispersiveGraphene.f90), which no reference driver calls, is applied
!     through its update_EJp in the same material pass: E form from curl H,
!     so those cells ignore D (test cases keep them apart from the others);
!     so is the E-form dielectric written for those simulations
!     (mod3/simpledielectricLorentzJp_mod.f90), mtype 11;
!   - the D-form Lorentz silicon (mod3/si_mod.f90) is mtype 12;
!   - source, DFT gating, probes and flux follow sim3/program.f90:396-685.
! Usage: ref3d <config.nml> <outdir>
program ref3d
    use constants_mod
    use mydatatypes
    use material_mod
    use simpleDielectric_mod
    use gold_mod
    use graphene_mod
    use BERpmly_mod
    use goldLorentzJp_mod
    use siLorentzJp_mod
    use titaniumJp_mod
    use filmgoldLorentzJp_mod
    use grapheneLorentzJp_mod
    use graphene3LorentzJp1nm_mod
    use nondispersivegrapheneJp_mod
    use simpledielectricLorentzJp_mod
    use si_mod
    implicit none
    integer :: nx, ny, nz, nsteps
    real*8 :: ds, Ku, dt
    logical :: pbcx, pbcz
    integer :: npmlx, npmly, npmlz
    real*8 :: order, R0, eps_low, eps_high, sigmam
    integer, parameter :: MAXM = 12
    integer :: nmat, mtype(MAXM), mbox(6, MAXM)
    real*8 :: meps(MAXM), gwNL(MAXM), gdnuNL(MAXM), gxi3(MAXM)
    logical :: s_on
    integer :: s_index, s_comp
    real*8 :: s_t0, s_st, s_w0, s_amp
    integer :: ndft, nfreq, d_index(3), d_nb(3), d_ne(3)
    real*8 :: lam_init, lam_end
    integer :: nprobe, p_ijk(3, 8)
    namelist /grid/ nx, ny, nz, ds, Ku, nsteps, pbcx, pbcz
    namelist /pml/ npmlx, npmly, npmlz, order, R0, eps_low, eps_high
    namelist /mats/ nmat, mtype, mbox, meps, gwNL, gdnuNL, gxi3
    namelist /src/ s_on, s_index, s_comp, s_t0, s_st, s_w0, s_amp
    namelist /dft/ ndft, nfreq, d_index, d_nb, d_ne, lam_init, lam_end
    namelist /probes/ nprobe, p_ijk

    real*8, allocatable, dimension(:,:,:), target :: Hx, Hy, Hz, Ex, Ey, Ez, Dx, Dy, Dz
    real*8, dimension(:,:,:), pointer :: PEx, PEy, PEz, PDx, PDy, PDz, PHx, PHy, PHz
    type(ConstantDielectric) :: diel(MAXM)
    type(Gold) :: au(MAXM)
    type(Graphene) :: gr(MAXM)
    ! the Lorentz-Jp family (mtype 4..10 = FDTD_LJP_* + 4), E form
    type(gold_LJp) :: lgold(MAXM)
    type(si_LJp) :: lsi(MAXM)
    type(Ti_LJp) :: lti(MAXM)
    type(filmgold_LJp) :: lfilm(MAXM)
    type(SLG_LJp) :: lgr(MAXM)
    type(SLG_3LJp1nm) :: lgr3(MAXM)
    type(SLG_ndJp) :: lgrnd(MAXM)
    ! mtype 11: the E-form dielectric of the Lorentz-Jp simulations
    ! (mod3/simpledielectricLorentzJp_mod.f90: E += Ku*ga*curl H)
    type(ConstantDielectricLJp) :: dlj(MAXM)
    ! mtype 12: the D-form Lorentz silicon (mod3/si_mod.f90)
    type(Si) :: sid(MAXM)
    type(pmly) :: lowp, highp
    type(Irange) :: absIx, absIy, absIz
    real*8, allocatable :: dre(:,:,:,:,:), dim_(:,:,:,:,:), arg(:), prb(:,:,:), Sy(:,:)
    real*8 :: t, gauss_t, freq, finit, fend, df, acc
    integer :: n, i, j, k, m, q, p, c, u
    character(len=512) :: cfg, outdir

    call get_command_argument(1, cfg)
    call get_command_argument(2, outdir)
    pbcx = .true.; pbcz = .true.
    npmlx = 10; npmly = 10; npmlz = 10; order = 3.d0; R0 = dexp(-16.d0)
    eps_low = 1.d0; eps_high = 1.d0
    nmat = 0; mtype = 0; mbox = 0; meps = 1.d0; gwNL = 0.d0; gdnuNL = 0.d0; gxi3 = 0.d0
    s_on = .false.; s_index = 1; s_comp = 1; s_t0 = 0.d0; s_st = 1.d0; s_w0 = 0.d0; s_amp = 1.d0
    ndft = 0; nfreq = 1; d_index = 1; d_nb = 1; d_ne = 0; lam_init = 500d-9; lam_end = 800d-9
    nprobe = 0; p_ijk = 1
    open(unit=7, file=trim(cfg), status='old')
    read(7, nml=grid)
    read(7, nml=pml)
    read(7, nml=mats)
    read(7, nml=src)
    read(7, nml=dft)
    read(7, nml=probes)
    close(7)

    dt = Ku*ds/c0
    allocate(Hx(0:nx+1,0:ny+1,0:nz+1), Hy(0:nx+1,0:ny+1,0:nz+1), Hz(0:nx+1,0:ny+1,0:nz+1))
    allocate(Ex(0:nx+1,0:ny+1,0:nz+1), Ey(0:nx+1,0:ny+1,0:nz+1), Ez(0:nx+1,0:ny+1,0:nz+1))
    allocate(Dx(0:nx+1,0:ny+1,0:nz+1), Dy(0:nx+1,0:ny+1,0:nz+1), Dz(0:nx+1,0:ny+1,0:nz+1))
    Hx = 0.d0; Hy = 0.d0; Hz = 0.d0; Ex = 0.d0; Ey = 0.d0; Ez = 0.d0
    Dx = 0.d0; Dy = 0.d0; Dz = 0.d0
    PEx => Ex; PEy => Ey; PEz => Ez
    PDx => Dx; PDy => Dy; PDz => Dz
    PHx => Hx; PHy => Hy; PHz => Hz

    sigmam = (order+1.d0)*c0*log(1.d0/R0)/(2.d0*npmly*ds)
    call lowp%initialize(eps_low*ep0, eps_high*ep0, sigmam, (/ npmlx, npmly, npmlz /), order, dt, ds, &
                         (/ 0, 0, 0 /), (/ nx, ny, nz /), (/ 1, 2, 1 /))
    call highp%initialize(eps_low*ep0, eps_high*ep0, sigmam, (/ npmlx, npmly, npmlz /), order, dt, ds, &
                          (/ 0, 1, 0 /), (/ nx, ny, nz /), (/ 1, 2, 1 /))

    absIx%init = 1; absIx%ends = nx
    absIy%init = 1; absIy%ends = ny
    absIz%init = 1; absIz%ends = nz
    do q = 1, nmat
        select case (mtype(q))
        case (1)
            call setbox(diel(q)%material, mbox(:, q))
            diel(q)%constant_ep = meps(q)
            call diel(q)%initialize_material(absIx, absIy, absIz, (/ nx, ny, nz /))
        case (2)
            call setbox(au(q)%material, mbox(:, q))
            au(q)%dt = dt
            call au(q)%initialize_material(absIx, absIy, absIz, (/ nx, ny, nz /))
        case (3)
            call setbox(gr(q)%material, mbox(:, q))
            gr(q)%dt = dt
            gr(q)%ds = ds
            gr(q)%wNL = gwNL(q)
            gr(q)%dnuNL = gdnuNL(q)
            gr(q)%xi3 = gxi3(q)
            call gr(q)%initialize_material(absIx, absIy, absIz, (/ nx, ny, nz /))
            print *, 'graphene gax(1st cell)=', gr(q)%getRgax(gr(q)%xrange%init, gr(q)%yrange%init, &
                gr(q)%zrange%init), gr(q)%getIgax(gr(q)%xrange%init, gr(q)%yrange%init, gr(q)%zrange%init)
        case (4)
            call setbox(lgold(q)%material, mbox(:, q))
            lgold(q)%dt = dt; lgold(q)%ds = ds
            call lgold(q)%initialize_material(absIx, absIy, absIz, (/ nx, ny, nz /))
        case (5)
            call setbox(lsi(q)%material, mbox(:, q))
            lsi(q)%dt = dt; lsi(q)%ds = ds
            call lsi(q)%initialize_material(absIx, absIy, absIz, (/ nx, ny, nz /))
        case (6)
            call setbox(lti(q)%material, mbox(:, q))
            lti(q)%dt = dt; lti(q)%ds = ds
            call lti(q)%initialize_material(absIx, absIy, absIz, (/ nx, ny, nz /))
        case (7)
            call setbox(lfilm(q)%material, mbox(:, q))
            lfilm(q)%dt = dt; lfilm(q)%ds = ds
            call lfilm(q)%initialize_material(absIx, absIy, absIz, (/ nx, ny, nz /))
        case (8)
            call setbox(lgr(q)%material, mbox(:, q))
            lgr(q)%dt = dt; lgr(q)%ds = ds
            lgr(q)%wNL = gwNL(q); lgr(q)%dnuNL = gdnuNL(q); lgr(q)%xi3 = gxi3(q)
            call lgr(q)%initialize_material(absIx, absIy, absIz, (/ nx, ny, nz /))
        case (9)
            call setbox(lgr3(q)%material, mbox(:, q))
            lgr3(q)%dt = dt; lgr3(q)%ds = ds
            lgr3(q)%wNL = gwNL(q); lgr3(q)%dnuNL = gdnuNL(q); lgr3(q)%xi3 = gxi3(q)
            call lgr3(q)%initialize_material(absIx, absIy, absIz, (/ nx, ny, nz /))
        case (10)
            call setbox(lgrnd(q)%material, mbox(:, q))
            lgrnd(q)%dt = dt; lgrnd(q)%ds = ds
            lgrnd(q)%wNL = gwNL(q); lgrnd(q)%dnuNL = gdnuNL(q); lgrnd(q)%xi3 = gxi3(q)
            call lgrnd(q)%initialize_material(absIx, absIy, absIz, (/ nx, ny, nz /))
        case (11)
            call setbox(dlj(q)%material, mbox(:, q))
            dlj(q)%dt = dt; dlj(q)%ds = ds
            dlj(q)%constant_ep = meps(q)
            call dlj(q)%initialize_material(absIx, absIy, absIz, (/ nx, ny, nz /))
        case (12)
            call setbox(sid(q)%material, mbox(:, q))
            sid(q)%dt = dt
            call sid(q)%initialize_material(absIx, absIy, absIz, (/ nx, ny, nz /))
        end select
    end do

    ! frequency grid, sim3/program.f90:121-125,353-357
    allocate(arg(nfreq))
    finit = c0/lam_end
    fend = c0/lam_init
    df = (fend-finit)/nfreq
    freq = finit
    do m = 1, nfreq
        arg(m) = 2.d0*pi*freq
        freq = freq+df
    end do
    allocate(dre(nx, nz, nfreq, 4, max(ndft, 1)), dim_(nx, nz, nfreq, 4, max(ndft, 1)))
    dre = 0.d0; dim_ = 0.d0
    allocate(prb(7, nsteps, max(nprobe, 1)))
    prb = 0.d0

    do n = 1, nsteps
        t = real(n-1, 8)*dt
        call updateD(1, nx, npmly+1, ny-npmly, 1, nz)
        if (s_on) then
            gauss_t = dexp(-0.5d0*((t-s_t0)/s_st)**2)*dsin(s_w0*t)
            do k = 1, nz
                do i = 1, nx
                    select case (s_comp)
                    case (1)
                        Dx(i, s_index, k) = Dx(i, s_index, k)+gauss_t*s_amp
                    case (2)
                        Dy(i, s_index, k) = Dy(i, s_index, k)+gauss_t*s_amp
                    case (3)
                        Dz(i, s_index, k) = Dz(i, s_index, k)+gauss_t*s_amp
                    end select
                end do
            end do
        end if
        do q = 1, nmat
            select case (mtype(q))
            case (1)
                call diel(q)%update_E(PEx, PEy, PEz, PDx, PDy, PDz)
            case (2)
                call au(q)%update_E(PEx, PEy, PEz, PDx, PDy, PDz)
            case (3)
                call gr(q)%update_E(PEx, PEy, PEz, PDx, PDy, PDz)
            case (4)
                call lgold(q)%update_EJp(PEx, PEy, PEz, PHx, PHy, PHz)
            case (5)
                call lsi(q)%update_EJp(PEx, PEy, PEz, PHx, PHy, PHz)
            case (6)
                call lti(q)%update_EJp(PEx, PEy, PEz, PHx, PHy, PHz)
            case (7)
                call lfilm(q)%update_EJp(PEx, PEy, PEz, PHx, PHy, PHz)
            case (8)
                call lgr(q)%update_EJp(PEx, PEy, PEz, PHx, PHy, PHz)
            case (9)
                call lgr3(q)%update_EJp(PEx, PEy, PEz, PHx, PHy, PHz)
            case (10)
                call lgrnd(q)%update_EJp(PEx, PEy, PEz, PHx, PHy, PHz)
            case (11)
                call dlj(q)%update_EJp(PEx, PEy, PEz, PHx, PHy, PHz)
            case (12)
                call sid(q)%update_E(PEx, PEy, PEz, PDx, PDy, PDz)
            end select
        end do
        call lowp%updateE(PEx, PEy, PEz, PHx, PHy, PHz)
        call highp%updateE(PEx, PEy, PEz, PHx, PHy, PHz)
        call ghost_E()
        call updateH(1, nx, npmly+1, ny-npmly, 1, nz)
        call lowp%updateH(PEx, PEy, PEz, PHx, PHy, PHz)
        call highp%updateH(PEx, PEy, PEz, PHx, PHy, PHz)
        call ghost_H()
        ! running DFT, sim3/program.f90:441-540 (y-normal planes)
        do p = 1, ndft
            if (n .lt. d_nb(p)) cycle
            if (d_ne(p) .gt. 0 .and. n .gt. d_ne(p)) cycle
            j = d_index(p)
            do m = 1, nfreq
                do k = 1, nz
                    do i = 1, nx
                        dre(i, k, m, 1, p) = dre(i, k, m, 1, p)+dcos(arg(m)*t)*ex(i, j, k)*dt
                        dim_(i, k, m, 1, p) = dim_(i, k, m, 1, p)+dsin(arg(m)*t)*ex(i, j, k)*dt
                        dre(i, k, m, 2, p) = dre(i, k, m, 2, p)+dcos(arg(m)*t)*ez(i, j, k)*dt
                        dim_(i, k, m, 2, p) = dim_(i, k, m, 2, p)+dsin(arg(m)*t)*ez(i, j, k)*dt
                        dre(i, k, m, 3, p) = dre(i, k, m, 3, p)+dcos(arg(m)*t)*hx(i, j, k)*dt
                        dim_(i, k, m, 3, p) = dim_(i, k, m, 3, p)+dsin(arg(m)*t)*hx(i, j, k)*dt
                        dre(i, k, m, 4, p) = dre(i, k, m, 4, p)+dcos(arg(m)*t)*hz(i, j, k)*dt
                        dim_(i, k, m, 4, p) = dim_(i, k, m, 4, p)+dsin(arg(m)*t)*hz(i, j, k)*dt
                    end do
                end do
            end do
        end do
        do p = 1, nprobe
            i = p_ijk(1, p); j = p_ijk(2, p); k = p_ijk(3, p)
            prb(1, n, p) = n*dt
            prb(2, n, p) = ex(i, j, k); prb(3, n, p) = ey(i, j, k); prb(4, n, p) = ez(i, j, k)
            prb(5, n, p) = hx(i, j, k); prb(6, n, p) = hy(i, j, k); prb(7, n, p) = hz(i, j, k)
        end do
    end do

    ! flux, sim3/program.f90:613-685
    allocate(Sy(nfreq, max(ndft, 1)))
    Sy = 0.d0
    do p = 1, ndft
        do m = 1, nfreq
            acc = 0.d0
            do k = 1, nz
                do i = 1, nx
                    acc = acc+( &
                        (dre(i, k, m, 3, p)*dre(i, k, m, 2, p))+ &
                        (dim_(i, k, m, 3, p)*dim_(i, k, m, 2, p))- &
                        (dre(i, k, m, 4, p)*dre(i, k, m, 1, p))- &
                        (dim_(i, k, m, 4, p)*dim_(i, k, m, 1, p)) &
                        )*ds
                end do
            end do
            Sy(m, p) = acc
        end do
    end do

    open(unit=20, file=trim(outdir)//'/fields.bin', access='stream', form='unformatted', status='replace')
    write(20) Ex(1:nx, 1:ny, 1:nz), Ey(1:nx, 1:ny, 1:nz), Ez(1:nx, 1:ny, 1:nz)
    write(20) Hx(1:nx, 1:ny, 1:nz), Hy(1:nx, 1:ny, 1:nz), Hz(1:nx, 1:ny, 1:nz)
    write(20) Dx(1:nx, 1:ny, 1:nz), Dy(1:nx, 1:ny, 1:nz), Dz(1:nx, 1:ny, 1:nz)
    close(20)
    open(unit=21, file=trim(outdir)//'/dft.bin', access='stream', form='unformatted', status='replace')
    write(21) arg
    do p = 1, ndft
        do c = 1, 4
            write(21) dre(:, :, :, c, p), dim_(:, :, :, c, p)
        end do
    end do
    write(21) Sy
    close(21)
    open(unit=22, file=trim(outdir)//'/probes.bin', access='stream', form='unformatted', status='replace')
    write(22) prb
    close(22)
    u = 0
    print *, 'done', nsteps, dt

contains
    include 'myfdtd_func.f90'

    subroutine setbox(mt, b)
        type(material), intent(inout) :: mt
        integer, intent(in) :: b(6)
        mt%xinit = b(1); mt%xend = b(2)
        mt%yinit = b(3); mt%yend = b(4)
        mt%zinit = b(5); mt%zend = b(6)
    end subroutine setbox

    ! interchange_E on one rank with Xmax = Zmax = 1 (mympi_func.f90:241-518)
    subroutine ghost_E()
        if (pbcx) then
            Ey(nx+1, :, :) = Ey(1, :, :)
            Ez(nx+1, :, :) = Ez(1, :, :)
        end if
        if (pbcz) then
            Ex(:, :, nz+1) = Ex(:, :, 1)
            Ey(:, :, nz+1) = Ey(:, :, 1)
        end if
        if (pbcx) then
            Ey(0, :, :) = Ey(nx, :, :)
            Ez(0, :, :) = Ez(nx, :, :)
        end if
        if (pbcz) then
            Ex(:, :, 0) = Ex(:, :, nz)
            Ey(:, :, 0) = Ey(:, :, nz)
        end if
    end subroutine ghost_E

    ! interchange_H on one rank with Xmax = Zmax = 1 (mympi_func.f90:522-690)
    subroutine ghost_H()
        if (pbcx) then
            Hy(0, :, :) = Hy(nx, :, :)
            Hz(0, :, :) = Hz(nx, :, :)
        end if
        if (pbcz) then
            Hx(:, :, 0) = Hx(:, :, nz)
            Hy(:, :, 0) = Hy(:, :, nz)
        end if
        if (pbcx) then
            Hy(nx+1, :, :) = Hy(1, :, :)
            Hz(nx+1, :, :) = Hz(1, :, :)
        end if
        if (pbcz) then
            Hx(:, :, nz+1) = Hx(:, :, 1)
            Hy(:, :, nz+1) = Hy(:, :, 1)
        end if
    end subroutine ghost_H
end program ref3d
