! Fortran binding of the gfx950 particle hot path (include/ipm_hip.h).
!
! The explicit bind(C) interfaces replace the reference's per-step particle
! calls; the wrappers at the bottom keep the reference's names and its error
! behaviour (print and stop, R/src/pic/particle/ParticleBundle.f90:587-590):
!
!   call PB(k)%MoveES(FG)                     -> call ipm_MoveES(ctx, k)
!   call MaterialAborptionParticle(..., PB(k)) -> call ipm_MaterialAborptionParticle(ctx, k, MT charge, sigma)
!   call FO(k)%P2CES(PB(k), Geom)             -> call ipm_P2CES(ctx, k, FO(k)%RhoOne, FO(k)%ChiOne)
!   call FT%Sum(FO)                           -> call ipm_FieldSourceSum(ctx, FT%Rho, FT%Chi)
!   call MCC(Ns, Ng, PB, SO, GO, MCCB)        -> call ipm_MCCStep(ctx)
!
! Species indices follow the reference (0 = electrons .. Ns); grid arrays are
! passed with the reference's own bounds and ghost layers (column-major).
module ipm_hip
    use iso_c_binding
    implicit none

    integer(c_int), parameter :: IPM_OK = 0, IPM_EINVAL = -1, IPM_EHIP = -2, IPM_ECAPACITY = -3, IPM_ENODEV = -4
    integer(c_int), parameter :: IPM_LAYOUT_AOS = 0, IPM_LAYOUT_SOA = 1

    interface
        integer(c_int) function ipm_create(ctx, nz, nr, dx, dt, ns, ng) bind(C, name="ipm_create")
            import :: c_ptr, c_int, c_double
            type(c_ptr), intent(out) :: ctx
            integer(c_int), value :: nz, nr, ns, ng
            real(c_double), value :: dx, dt
        end function
        integer(c_int) function ipm_destroy(ctx) bind(C, name="ipm_destroy")
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
        end function
        type(c_ptr) function ipm_last_error() bind(C, name="ipm_last_error")
            import :: c_ptr
        end function
        integer(c_int) function ipm_set_species(ctx, s, charge, mass, capacity) bind(C, name="ipm_set_species")
            import :: c_ptr, c_int, c_double, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int), value :: s
            real(c_double), value :: charge, mass
            integer(c_int64_t), value :: capacity
        end function
        integer(c_int) function ipm_set_geometry(ctx, inv_volume, cell_material, hedge_type, vedge_type, &
                                                 ndiel, nmetal, material_type) bind(C, name="ipm_set_geometry")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: inv_volume(*)
            integer(c_int), intent(in) :: cell_material(*), hedge_type(*), vedge_type(*), material_type(*)
            integer(c_int), value :: ndiel, nmetal
        end function
        integer(c_int) function ipm_upload_species(ctx, s, n, po, layout) bind(C, name="ipm_upload_species")
            import :: c_ptr, c_int, c_double, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int), value :: s, layout
            integer(c_int64_t), value :: n
            real(c_double), intent(in) :: po(*)
        end function
        integer(c_int) function ipm_download_species(ctx, s, n, po, cap, layout) bind(C, name="ipm_download_species")
            import :: c_ptr, c_int, c_double, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int), value :: s, layout
            integer(c_int64_t), intent(out) :: n
            real(c_double), intent(inout) :: po(*)
            integer(c_int64_t), value :: cap
        end function
        integer(c_int) function ipm_count(ctx, s, n) bind(C, name="ipm_count")
            import :: c_ptr, c_int, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int), value :: s
            integer(c_int64_t), intent(out) :: n
        end function
        integer(c_int) function ipm_set_fields(ctx, Ez, Er) bind(C, name="ipm_set_fields")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: Ez(*), Er(*)
        end function
        integer(c_int) function ipm_move_es(ctx, s) bind(C, name="ipm_move_es")
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int), value :: s
        end function
        integer(c_int) function ipm_absorb(ctx, s) bind(C, name="ipm_absorb")
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int), value :: s
        end function
        integer(c_int) function ipm_move_absorb(ctx, s) bind(C, name="ipm_move_absorb")
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int), value :: s
        end function
        integer(c_int) function ipm_take_tallies(ctx, metal_charge, sigma) bind(C, name="ipm_take_tallies")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(inout) :: metal_charge(*), sigma(*)
        end function
        integer(c_int) function ipm_p2c(ctx, s, rho_one, chi_one) bind(C, name="ipm_p2c")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: s
            real(c_double), intent(inout) :: rho_one(*), chi_one(*)
        end function
        integer(c_int) function ipm_field_sum(ctx, rho, chi) bind(C, name="ipm_field_sum")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(inout) :: rho(*), chi(*)
        end function
        integer(c_int) function ipm_set_mcc(ctx, s, g, model, nreaction, nsigma, emin, einterval, emax, &
                                            collision_ratio, rtype, resultant, threshold, prob, gas_mass, &
                                            gas_temperature, beta_max) bind(C, name="ipm_set_mcc")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: s, g, model, nreaction, nsigma
            real(c_double), value :: emin, einterval, emax, collision_ratio, gas_mass, gas_temperature, beta_max
            integer(c_int), intent(in) :: rtype(*), resultant(*)
            real(c_double), intent(in) :: threshold(*), prob(*)
        end function
        integer(c_int) function ipm_set_rng(ctx, seed, step) bind(C, name="ipm_set_rng")
            import :: c_ptr, c_int, c_int64_t, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: seed
            integer(c_int32_t), value :: step
        end function
        integer(c_int) function ipm_mcc(ctx) bind(C, name="ipm_mcc")
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function ipm_set_event_counting(ctx, on) bind(C, name="ipm_set_event_counting")
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int), value :: on
        end function
        integer(c_int) function ipm_set_null_screen(ctx, on) bind(C, name="ipm_set_null_screen")
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int), value :: on
        end function
        integer(c_int) function ipm_take_mcc_counts(ctx, counts) bind(C, name="ipm_take_mcc_counts")
            import :: c_ptr, c_int, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int64_t), intent(inout) :: counts(*)
        end function
        integer(c_int) function ipm_step(ctx) bind(C, name="ipm_step")
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function ipm_step_graph(ctx) bind(C, name="ipm_step_graph")
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function ipm_set_timing(ctx, on) bind(C, name="ipm_set_timing")
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int), value :: on
        end function
        integer(c_int) function ipm_take_timing(ctx, ms, launches) bind(C, name="ipm_take_timing")
            import :: c_ptr, c_int, c_double, c_int64_t
            type(c_ptr), value :: ctx
            real(c_double), intent(inout) :: ms(*)
            integer(c_int64_t), intent(inout) :: launches(*)
        end function
        integer(c_int) function ipm_sort(ctx, s) bind(C, name="ipm_sort")
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int), value :: s
        end function
        integer(c_int) function ipm_set_launch(ctx, block, blocks_per_cu) bind(C, name="ipm_set_launch")
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int), value :: block, blocks_per_cu
        end function
        integer(c_int) function ipm_set_collide_grid(ctx, blocks_per_cu) bind(C, name="ipm_set_collide_grid")
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int), value :: blocks_per_cu
        end function
        integer(c_int) function ipm_set_deposit_window(ctx, rows) bind(C, name="ipm_set_deposit_window")
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int), value :: rows
        end function
        integer(c_int) function ipm_set_sort_tile(ctx, slots) bind(C, name="ipm_set_sort_tile")
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int), value :: slots
        end function
        integer(c_int) function ipm_set_sort_interval(ctx, steps) bind(C, name="ipm_set_sort_interval")
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int), value :: steps
        end function
        integer(c_int) function ipm_set_step_rezone(ctx, s, mode, weight, alpha, grid_r, merge_limit) &
                bind(C, name="ipm_set_step_rezone")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: s, mode
            real(c_double), value :: weight, alpha, grid_r, merge_limit
        end function
        integer(c_int) function ipm_upload_species_device(ctx, s, offset, n, po_dev, ld) &
                bind(C, name="ipm_upload_species_device")
            import :: c_ptr, c_int, c_int64_t
            type(c_ptr), value :: ctx, po_dev
            integer(c_int), value :: s
            integer(c_int64_t), value :: offset, n, ld
        end function
        integer(c_int) function ipm_moments_noalloc(ctx, s, implicit_j, rho, energy_ev, mp, jz, jr) &
                bind(C, name="ipm_moments_noalloc")
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx, rho, energy_ev, mp, jz, jr
            integer(c_int), value :: s, implicit_j
        end function
        integer(c_int) function ipm_window_info(ctx, nparticles, nsegments, forward_ms, adjoint_ms) &
                bind(C, name="ipm_window_info")
            import :: c_ptr, c_int, c_int64_t, c_float
            type(c_ptr), value :: ctx
            integer(c_int64_t), intent(out) :: nparticles
            integer(c_int), intent(out) :: nsegments
            real(c_float), intent(out) :: forward_ms, adjoint_ms
        end function
        integer(c_int) function ipm_window_download(ctx, checkpoint, state, absorbed_step) &
                bind(C, name="ipm_window_download")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: checkpoint
            real(c_double), intent(inout) :: state(*)
            integer(c_int), intent(inout) :: absorbed_step(*)
        end function
        integer(c_int) function ipm_synchronize(ctx) bind(C, name="ipm_synchronize")
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function ipm_species_device_ptrs(ctx, s, fields8) bind(C, name="ipm_species_device_ptrs")
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int), value :: s
            type(c_ptr), intent(inout) :: fields8(8)
        end function
        integer(c_int) function ipm_set_potential(ctx, phi) bind(C, name="ipm_set_potential")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: phi(*)
        end function
        integer(c_int) function ipm_set_solver_geometry(ctx, cell_epsilon, point_type) &
                bind(C, name="ipm_set_solver_geometry")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: cell_epsilon(*)
            integer(c_int), intent(in) :: point_type(*)
        end function
        integer(c_int) function ipm_solver_system(ctx, with_chi, sigma, bound, bound_mask, coe, source) &
                bind(C, name="ipm_solver_system")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: with_chi
            real(c_double), intent(in) :: sigma(*), bound(*)
            integer(c_int), intent(in) :: bound_mask(*)
            real(c_double), intent(inout) :: coe(*), source(*)
        end function
        integer(c_int) function ipm_moments(ctx, s, implicit_j, rho, energy_ev, mp, jz, jr) &
                bind(C, name="ipm_moments")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: s, implicit_j
            real(c_double), intent(inout) :: rho(*), energy_ev(*), mp(*), jz(*), jr(*)
        end function
        integer(c_int) function ipm_recombine(ctx, nreaction, reaction_list, rate, time_step, deleted) &
                bind(C, name="ipm_recombine")
            import :: c_ptr, c_int, c_double, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int), value :: nreaction
            integer(c_int), intent(in) :: reaction_list(*)
            real(c_double), intent(in) :: rate(*)
            real(c_double), value :: time_step
            integer(c_int64_t), intent(inout) :: deleted(*)
        end function
        integer(c_int) function ipm_split_merge(ctx, s, mode, weight, alpha, grid_r, merge_limit, stats) &
                bind(C, name="ipm_split_merge")
            import :: c_ptr, c_int, c_double, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int), value :: s, mode
            real(c_double), value :: weight, alpha, grid_r, merge_limit
            integer(c_int64_t), intent(inout) :: stats(*)
        end function
        integer(c_int) function ipm_window_create(ctx, nsteps, interval, Ez0, Er0, Ez1, Er1, mu, alpha) &
                bind(C, name="ipm_window_create")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nsteps, interval
            real(c_double), intent(in) :: Ez0(*), Er0(*), Ez1(*), Er1(*), mu(*), alpha(*)
        end function
        integer(c_int) function ipm_window_forward(ctx, coef, record, loss) bind(C, name="ipm_window_forward")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: coef(*)
            integer(c_int), value :: record
            real(c_double), intent(inout) :: loss(*)
        end function
        integer(c_int) function ipm_window_backward(ctx, coef, gloss, dcoef) bind(C, name="ipm_window_backward")
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: coef(*), gloss(*)
            real(c_double), intent(inout) :: dcoef(*)
        end function
        integer(c_int) function ipm_window_destroy(ctx) bind(C, name="ipm_window_destroy")
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
        end function
        integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
            import :: c_ptr, c_size_t
            type(c_ptr), value :: s
        end function
    end interface

contains

    ! Text of the library's last error.
    function ipm_error_text() result(txt)
        character(len=:), allocatable :: txt
        character(kind=c_char), pointer :: chars(:)
        type(c_ptr) :: p
        integer :: n, i
        p = ipm_last_error()
        n = int(c_strlen(p))
        call c_f_pointer(p, chars, [n])
        allocate(character(len=n) :: txt)
        do i = 1, n
            txt(i:i) = chars(i)
        end do
    end function

    ! The reference prints and stops on every hot-path error; so does this.
    subroutine ipm_check(rc, where)
        integer(c_int), intent(in) :: rc
        character(*), intent(in) :: where
        if (rc /= IPM_OK) then
            write(*, '(a, a, a, i0, a, a)') "ipm: ", where, " failed (", rc, "): ", ipm_error_text()
            stop 1
        end if
    end subroutine

    ! PB(k)%MoveES(FG): R/src/pic/particle/ParticleBundle.f90:649
    subroutine ipm_MoveES(ctx, k)
        type(c_ptr), intent(in) :: ctx
        integer, intent(in) :: k
        call ipm_check(ipm_move_es(ctx, int(k, c_int)), "MoveES")
    end subroutine

    ! MaterialAborptionParticle(CF, Geom, MT, PB(k)):
    ! R/src/pic/particleboundary/ParticleMaterial.F90:10.  charge_one_step(1:nmetal)
    ! and sigma_one_step(1:Nz, 1:Nr) are accumulated, as the reference does.
    subroutine ipm_MaterialAborptionParticle(ctx, k, charge_one_step, sigma_one_step)
        type(c_ptr), intent(in) :: ctx
        integer, intent(in) :: k
        real(c_double), intent(inout) :: charge_one_step(:), sigma_one_step(:, :)
        call ipm_check(ipm_absorb(ctx, int(k, c_int)), "MaterialAborptionParticle")
        call ipm_check(ipm_take_tallies(ctx, charge_one_step, sigma_one_step), "MaterialAborptionParticle")
    end subroutine

    ! FO(k)%P2CES(PB(k), Geom): R/src/pic/field/FieldOne.f90:132
    subroutine ipm_P2CES(ctx, k, RhoOne, ChiOne)
        type(c_ptr), intent(in) :: ctx
        integer, intent(in) :: k
        real(c_double), intent(inout) :: RhoOne(:, :), ChiOne(:, :)
        call ipm_check(ipm_p2c(ctx, int(k, c_int), RhoOne, ChiOne), "P2CES")
    end subroutine

    ! FT%Sum(FO): R/src/pic/field/FieldSource.f90:72.  Rho(1:Nz, 1:Nr),
    ! Chi(0:Nz+1, 0:Nr+1) as the reference allocates them.
    subroutine ipm_FieldSourceSum(ctx, Rho, Chi)
        type(c_ptr), intent(in) :: ctx
        real(c_double), intent(inout) :: Rho(:, :), Chi(:, :)
        call ipm_check(ipm_field_sum(ctx, Rho, Chi), "FieldSource%Sum")
    end subroutine

    ! MCC(Ns, Ng, PB, SO, GO, MCCB): R/src/mc/MCCPublic.F90:38
    subroutine ipm_MCCStep(ctx)
        type(c_ptr), intent(in) :: ctx
        call ipm_check(ipm_mcc(ctx), "MCC")
    end subroutine

    ! FT%SetE(FG): R/src/pic/field/FieldSource.f90:95, from Phi(1:Nz, 1:Nr).
    subroutine ipm_SetE(ctx, Phi)
        type(c_ptr), intent(in) :: ctx
        real(c_double), intent(in) :: Phi(:, :)
        call ipm_check(ipm_set_potential(ctx, Phi), "SetE")
    end subroutine

    ! FS%SetCoe(FT, Geom), FS%SetSrc(FT, MT, Geom) and FS%SetBound(boundary)
    ! (R/src/pic/field/FieldSolver.f90:136-248) assembled on the device from
    ! the Rho and Chi of the last FieldSourceSum.  sigma = MT%sigma, bound and
    ! bound_mask = boundary%bound / %bound_mask; every array is (1:Nz, 1:Nr).
    ! Geom%cell_epsilon and Geom%point_type go down once, through
    ! ipm_set_solver_geometry.
    subroutine ipm_SetSolverSystem(ctx, sigma, bound, bound_mask, CoeA, CoeB, CoeC, CoeD, CoeE, Source)
        type(c_ptr), intent(in) :: ctx
        real(c_double), intent(in) :: sigma(:, :), bound(:, :)
        integer(c_int), intent(in) :: bound_mask(:, :)
        real(c_double), intent(inout) :: CoeA(:, :), CoeB(:, :), CoeC(:, :), CoeD(:, :), CoeE(:, :), Source(:, :)
        real(c_double), allocatable :: coe(:, :, :)
        allocate(coe(size(CoeA, 1), size(CoeA, 2), 5))
        call ipm_check(ipm_solver_system(ctx, 1_c_int, sigma, bound, bound_mask, coe, Source), "SetCoe")
        CoeA = coe(:, :, 1)
        CoeB = coe(:, :, 2)
        CoeC = coe(:, :, 3)
        CoeD = coe(:, :, 4)
        CoeE = coe(:, :, 5)
        deallocate(coe)
    end subroutine

    ! The electrons' every-step split_particle_weight + merge_particle_random
    ! (particle_split_merge_control, R/src/base/Rezoning.F90:33-34) folded into
    ! every later ipm_step / ipm_step_graph; call again whenever PB(k)%Weight
    ! changes (every CF%Period).
    subroutine ipm_SplitMergeInStep(ctx, k, Weight, alpha, grid_r, merge_limit)
        type(c_ptr), intent(in) :: ctx
        integer, intent(in) :: k
        real(c_double), intent(in) :: Weight, alpha, grid_r, merge_limit
        call ipm_check(ipm_set_step_rezone(ctx, int(k, c_int), 3_c_int, Weight, alpha, grid_r, merge_limit), &
                       "SplitMergeInStep")
    end subroutine

    ! WeightingParticleField for species k: R/src/diag/diagParticleField.F90:173.
    ! RhoOne, TOne (eV), MPOne, JzOne, JrOne of DiagOne(k), each (1:Nz, 1:Nr);
    ! the current from a copy pushed once more, as for the implicit scheme.
    subroutine ipm_WeightingParticleField(ctx, k, RhoOne, TOne, MPOne, JzOne, JrOne)
        type(c_ptr), intent(in) :: ctx
        integer, intent(in) :: k
        real(c_double), intent(inout) :: RhoOne(:, :), TOne(:, :), MPOne(:, :), JzOne(:, :), JrOne(:, :)
        call ipm_check(ipm_moments(ctx, int(k, c_int), 1_c_int, RhoOne, TOne, MPOne, JzOne, JrOne), &
                       "WeightingParticleField")
    end subroutine

    ! split_particle_weight then merge_particle_random of ModuleParticleRezoning
    ! (R/src/base/Rezoning.F90:53-92) for bundle k; Weight, alpha and grid_r are
    ! PB(k)%Weight, PB(k)%alpha and PB(k)%grid_r, merge_limit the module's.
    subroutine ipm_SplitMerge(ctx, k, Weight, alpha, grid_r, merge_limit)
        type(c_ptr), intent(in) :: ctx
        integer, intent(in) :: k
        real(c_double), intent(in) :: Weight, alpha, grid_r, merge_limit
        integer(c_int64_t) :: stats(3)
        call ipm_check(ipm_split_merge(ctx, int(k, c_int), 3_c_int, Weight, alpha, grid_r, merge_limit, stats), &
                       "SplitMerge")
    end subroutine

    ! RecombinationLocal%recb(Ns, PB, time_step) (R/src/mc/Recombination.F90:164-243)
    ! with Re2d%recb_reaction(reaction_num, 3) (0-based species, as stored) and
    ! Re2d%recb_rate; time_step = recombination_time_step_number * dt.
    subroutine ipm_Recombination(ctx, ns, recb_reaction, recb_rate, time_step)
        type(c_ptr), intent(in) :: ctx
        integer, intent(in) :: ns
        integer(c_int), intent(in) :: recb_reaction(:, :)
        real(c_double), intent(in) :: recb_rate(:), time_step
        integer(c_int64_t) :: deleted(ns + 1)
        call ipm_check(ipm_recombine(ctx, int(size(recb_reaction, 1), c_int), recb_reaction, recb_rate, time_step, &
                                     deleted), "Recombination")
    end subroutine

    ! PB(k)%NPar
    integer(c_int64_t) function ipm_NPar(ctx, k)
        type(c_ptr), intent(in) :: ctx
        integer, intent(in) :: k
        call ipm_check(ipm_count(ctx, int(k, c_int), ipm_NPar), "NPar")
    end function

end module ipm_hip
