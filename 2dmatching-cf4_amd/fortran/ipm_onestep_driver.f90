! Fortran host driver: the reference's OneStep particle sequence
! (R/src/onestep.F90:141-177) on the GPU through the ipm_hip module.
!
!   ipm_onestep_driver <outdir> [steps]
!
! Builds a small capacitive cylinder (33 x 17 nodes, electrodes of metal
! cells at both ends), loads electrons and CF3+ ions uniformly in volume
! (Wq proportional to the cell's radial index, as AllInit does), sets a
! synthetic field from an RF potential with the SetE stencil, builds a
! constant-sigma elastic electron table the way ProbilityNonReactive does,
! then:
!   1. writes the field (Ez then Er) to <outdir>/fields.bin and species 0
!      (AoS, ParticleOne layout) to <outdir>/pre.bin,
!      calls MoveES for species 0 only and writes the result to post.bin;
!   2. runs [steps] full steps (MoveES + absorption per species, P2CES per
!      species, FieldSource%Sum, MCC) and prints NPar, the metal charge and
!      the deposited charge against the charge carried by the particles;
!   3. after the last step's FieldSource%Sum, assembles the field solver's
!      system (SetCoe, SetSrc, SetBound) on the device and writes Rho, Chi,
!      CoeA..CoeE and Source to <outdir>/system.bin.
program ipm_onestep_driver
    use iso_c_binding
    use ipm_hip
    implicit none

    integer, parameter :: Nz = 33, Nr = 17, Ns = 2, NSIG = 10001
    real(8), parameter :: PI = 3.141592653589793238d0, QE = 1.6022d-19, ME = 9.1095d-31
    real(8), parameter :: AMU = 1.660539040d-27, KB = 1.3807d-23, JtoeV = 1.6022d-19
    real(8), parameter :: dx = 1.d-3, dt = 5.d-11
    type(c_ptr) :: ctx
    character(len=256) :: outdir, arg
    integer :: nsteps, i, j, k, s, ppc, n, it
    real(8) :: charge(0:Ns-1), mass(0:Ns-1), temp(0:Ns-1)
    integer(c_int), allocatable :: cell_material(:, :), hedge(:, :), vedge(:, :), mtype(:)
    real(8), allocatable :: vfac(:, :), vol(:, :), inv_vol(:, :)
    real(8), allocatable :: PO(:, :), phi(:, :), Ez(:, :), Er(:, :)
    real(8), allocatable :: RhoOne(:, :), ChiOne(:, :), Rho(:, :), Chi(:, :), sigma(:, :)
    real(8), allocatable :: eps(:, :), bound(:, :), CoeA(:, :), CoeB(:, :), CoeC(:, :), CoeD(:, :), CoeE(:, :)
    real(8), allocatable :: Source(:, :)
    integer(c_int), allocatable :: point_type(:, :), bound_mask(:, :)
    real(8) :: metal_charge(2), wq, r, z, u1, u2, vth, vf, q_particles, q_grid, q_abs
    integer(c_int64_t) :: nget, npar
    integer(c_int) :: rt(1), res(1)
    real(8) :: th(1), emin, eint, energy, v, pmax, ngas, sigma0, ratio
    real(8), allocatable :: prob(:)

    call get_command_argument(1, outdir)
    if (len_trim(outdir) == 0) outdir = "."
    nsteps = 5
    call get_command_argument(2, arg)
    if (len_trim(arg) > 0) read(arg, *) nsteps

    charge = [-QE, QE]
    mass = [(ME / AMU) * AMU, 69.0059096d0 * AMU]
    temp = [3.d0 * 11605.d0, 300.d0]

    call ipm_check(ipm_create(ctx, Nz, Nr, dx, dt, Ns, 1), "create")
    do s = 0, Ns - 1
        call ipm_check(ipm_set_species(ctx, s, charge(s), mass(s), int(400000, c_int64_t)), "set_species")
    end do

    ! geometry: metal 1 (powered) cells iz = 1, metal 2 (grounded) iz = Nz-1
    allocate(cell_material(Nz-1, Nr-1), hedge(Nz, Nr-1), vedge(Nz-1, Nr), mtype(0:2))
    cell_material = 0
    cell_material(1, :) = 1
    cell_material(Nz-1, :) = 2
    mtype = [0, 1, 1]                       ! vacuum, metal, metal (no dielectrics)
    hedge = 4
    vedge = 4
    hedge(2, :) = 2                         ! metal -> vacuum
    hedge(Nz-1, :) = 0                      ! vacuum -> metal
    ! volume factors as the geometry generator's update_volume gives them for
    ! this layout: the fraction of vacuum among a node's neighbouring cells in
    ! z (the radial shape is in the volume formula below)
    allocate(vfac(Nz, Nr), vol(Nz, Nr), inv_vol(Nz, Nr))
    vfac = 1.d0
    vfac(1, :) = 0.d0
    vfac(2, :) = 0.5d0
    vfac(Nz-1, :) = 0.5d0
    vfac(Nz, :) = 0.d0
    do j = 1, Nr
        do i = 1, Nz
            if (j == 1) then
                vol(i, j) = vfac(i, j) * 1.d0 / 3.d0 * PI * dx * dx * dx
            else if (j == Nr) then
                vol(i, j) = vfac(i, j) * 1.d0 / 3.d0 * (dble(j-2) + 2.d0 * dble(j-1)) * PI * dx * dx * dx
            else
                vol(i, j) = vfac(i, j) * 1.d0 / 3.d0 * (dble(j) * (dble(j-1) + dble(j)) - &
                            dble(j-2) * (dble(j-2) + dble(j-1))) * PI * dx * dx * dx
            end if
            inv_vol(i, j) = 0.d0
            if (vfac(i, j) > 1.d-6) inv_vol(i, j) = 1.d0 / vol(i, j)
        end do
    end do
    call ipm_check(ipm_set_geometry(ctx, inv_vol, cell_material, hedge, vedge, 0, 2, mtype), "set_geometry")
    ! field-solve geometry: relative permittivity growing with the cell
    ! indices (distinct per cell), metal nodes on the electrode faces; the
    ! powered electrode held at 100 V, the grounded one at 0
    allocate(eps(Nz-1, Nr-1), point_type(Nz, Nr), bound(Nz, Nr), bound_mask(Nz, Nr))
    do j = 1, Nr - 1
        do i = 1, Nz - 1
            eps(i, j) = 1.d0 + 0.01d0 * dble(i) + 0.1d0 * dble(j)
        end do
    end do
    point_type = 0
    point_type(1:2, :) = 1
    point_type(Nz-1:Nz, :) = 2
    bound = 0.d0
    bound(1:2, :) = 100.d0
    bound_mask = merge(1, 0, point_type > 0)
    call ipm_check(ipm_set_solver_geometry(ctx, eps, point_type), "set_solver_geometry")

    ! electron elastic table (Model 1, one reaction, constant sigma) as
    ! ProbilityNonReactive builds it: P = sigma v n / max over the log grid
    emin = -3.d0
    eint = 6.d0 / dble(NSIG - 1)
    ngas = (50.d0 * 0.13332d0) / (KB * 300.d0)
    sigma0 = 1.d-19
    allocate(prob(NSIG))
    do i = 1, NSIG
        energy = 10.d0**(emin + dble(i-1) * eint)
        v = dsqrt(2.d0 * energy * JtoeV / (mass(0) * 88.0043128d0 * AMU / (mass(0) + 88.0043128d0 * AMU)))
        prob(i) = sigma0 * v * ngas
    end do
    pmax = maxval(prob)
    prob = prob / pmax
    ratio = 1.d0 - dexp(-pmax * dt)
    rt = [11]
    res = [0]
    th = [0.d0]
    call ipm_check(ipm_set_mcc(ctx, 0, 0, 1, 1, NSIG, emin, eint, 1.d3, ratio, rt, res, th, prob, &
                               88.0043128d0 * AMU, 300.d0, 3.d0), "set_mcc")
    call ipm_check(ipm_set_mcc(ctx, 1, 0, 0, 0, 0, 0.d0, 0.d0, 0.d0, 0.d0, rt, res, th, prob, 0.d0, 0.d0, 0.d0), &
                   "set_mcc")
    call ipm_check(ipm_set_rng(ctx, int(12345, c_int64_t), int(0, c_int32_t)), "set_rng")

    ! particles: ppc per vacuum cell, uniform in volume, Maxwellian
    call random_seed()
    ppc = 16
    vf = dx / dt
    q_particles = 0.d0
    do s = 0, Ns - 1
        allocate(PO(8, ppc * (Nz-3) * (Nr-1)))
        n = 0
        vth = dsqrt(KB * temp(s) / mass(s))
        do i = 2, Nz - 2
            do j = 1, Nr - 1
                wq = 1.d15 * PI * dx**3 * (2.d0 * dble(j) - 1.d0) / dble(ppc)
                do k = 1, ppc
                    n = n + 1
                    call random_number(u1)
                    call random_number(u2)
                    z = dble(i - 1) + u1
                    r = dsqrt(u2 * (dble(j)**2 - dble(j-1)**2) + dble(j-1)**2)
                    PO(1, n) = z
                    PO(2, n) = max(r, dble(j-1) + 1.d-9)
                    do it = 3, 5
                        call random_number(u1)
                        call random_number(u2)
                        PO(it, n) = vth / vf * dsqrt(-2.d0 * dlog(max(u1, 1.d-300))) * dcos(2.d0 * PI * u2)
                    end do
                    PO(6:7, n) = 0.d0
                    PO(8, n) = wq
                end do
            end do
        end do
        call ipm_check(ipm_upload_species(ctx, s, int(n, c_int64_t), PO, IPM_LAYOUT_AOS), "upload")
        deallocate(PO)
    end do

    ! field: RF potential on z = 0 falling to ground at z = L plus a plasma
    ! bump, SetE with the linear ghost extrapolation
    allocate(phi(0:Nz+1, 0:Nr+1), Ez(0:Nz, 1:Nr), Er(1:Nz, 0:Nr))
    do j = 1, Nr
        do i = 1, Nz
            phi(i, j) = 150.d0 * (1.d0 - dble(i-1) / dble(Nz-1)) + &
                        20.d0 * dcos(PI * dble(j-1) / dble(2*(Nr-1)))**2 * dsin(PI * dble(i-1) / dble(Nz-1))**2
        end do
    end do
    phi(0, 1:Nr) = 2.d0 * phi(1, 1:Nr) - phi(2, 1:Nr)
    phi(Nz+1, 1:Nr) = 2.d0 * phi(Nz, 1:Nr) - phi(Nz-1, 1:Nr)
    phi(1:Nz, 0) = 2.d0 * phi(1:Nz, 1) - phi(1:Nz, 2)
    phi(1:Nz, Nr+1) = 2.d0 * phi(1:Nz, Nr) - phi(1:Nz, Nr-1)
    do i = 0, Nz
        Ez(i, :) = (phi(i, 1:Nr) - phi(i+1, 1:Nr)) / dx
    end do
    do j = 0, Nr
        Er(:, j) = (phi(1:Nz, j) - phi(1:Nz, j+1)) / dx
    end do
    call ipm_check(ipm_set_fields(ctx, Ez, Er), "set_fields")
    open(10, file=trim(outdir)//"/fields.bin", access="stream", form="unformatted", status="replace")
    write(10) Ez, Er
    close(10)

    ! 1. one MoveES of species 0, before and after
    npar = ipm_NPar(ctx, 0)
    allocate(PO(8, npar))
    call ipm_check(ipm_download_species(ctx, 0, nget, PO, npar, IPM_LAYOUT_AOS), "download")
    open(10, file=trim(outdir)//"/pre.bin", access="stream", form="unformatted", status="replace")
    write(10) nget, PO
    close(10)
    call ipm_MoveES(ctx, 0)
    call ipm_check(ipm_download_species(ctx, 0, nget, PO, npar, IPM_LAYOUT_AOS), "download")
    open(10, file=trim(outdir)//"/post.bin", access="stream", form="unformatted", status="replace")
    write(10) nget, PO
    close(10)
    deallocate(PO)
    call ipm_check(ipm_move_es(ctx, 1), "MoveES")      ! the ions of the same step

    ! 2. full steps in the reference's order
    allocate(RhoOne(Nz, Nr), ChiOne(Nz, Nr), Rho(Nz, Nr), Chi(0:Nz+1, 0:Nr+1), sigma(Nz, Nr))
    metal_charge = 0.d0
    sigma = 0.d0
    do it = 1, nsteps
        if (it > 1) then
            do s = 0, Ns - 1
                call ipm_MoveES(ctx, s)
            end do
        end if
        do s = 0, Ns - 1
            call ipm_MaterialAborptionParticle(ctx, s, metal_charge, sigma)
        end do
        do s = 0, Ns - 1
            call ipm_P2CES(ctx, s, RhoOne, ChiOne)
        end do
        call ipm_FieldSourceSum(ctx, Rho, Chi)
        ! deposited charge against the particles' charge (interior nodes keep
        ! all of it: no particle sits in the zero-volume end layers)
        q_grid = sum(Rho * vol)
        q_particles = 0.d0
        q_abs = 0.d0
        do s = 0, Ns - 1
            npar = ipm_NPar(ctx, s)
            allocate(PO(8, max(npar, 1_c_int64_t)))
            call ipm_check(ipm_download_species(ctx, s, nget, PO, max(npar, 1_c_int64_t), IPM_LAYOUT_AOS), "download")
            q_particles = q_particles + charge(s) * sum(PO(8, 1:nget))
            q_abs = q_abs + abs(charge(s)) * sum(PO(8, 1:nget))
            deallocate(PO)
        end do
        if (it == nsteps) then
            allocate(CoeA(Nz, Nr), CoeB(Nz, Nr), CoeC(Nz, Nr), CoeD(Nz, Nr), CoeE(Nz, Nr), Source(Nz, Nr))
            call ipm_SetSolverSystem(ctx, sigma, bound, bound_mask, CoeA, CoeB, CoeC, CoeD, CoeE, Source)
            open(newunit=k, file=trim(outdir)//"/system.bin", access="stream", form="unformatted", &
                 status="replace")
            write(k) Rho, Chi, sigma, inv_vol, CoeA, CoeB, CoeC, CoeD, CoeE, Source
            close(k)
        end if
        call ipm_MCCStep(ctx)
        write(*, '(a, i3, a, i8, a, i8, a, es24.16, a, es24.16, a, es24.16, a, 2es12.4)') "step", it, &
            " NPar(e)", ipm_NPar(ctx, 0), " NPar(CF3+)", ipm_NPar(ctx, 1), &
            " q_grid", q_grid, " q_particles", q_particles, " q_abs", q_abs, " metal", metal_charge
    end do
    call ipm_check(ipm_destroy(ctx), "destroy")
    write(*, '(a)') "driver done"
end program ipm_onestep_driver
