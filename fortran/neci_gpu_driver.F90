! neci_gpu_driver -- FciMCPar-style host loop in Fortran over the GPU engine.
!
! Real-space Hubbard chain (SYSTEM HUBBARD REAL-SPACE / LATTICE CHAIN n 1),
! Neel reference (create_neel_state_chain), plain or initiator FCIQMC.  The
! update cycle follows update_shift (fcimc_iter_utilities.F90:883-1260):
! single-particle phase until TotParts > TOTALWALKERS, jump to the projected
! energy, then DiagSft -= ln(growth) * SftDamp / (tau * StepsSft).
!
! usage: neci_gpu_driver nsites nel U tau niter totalwalkers seed truncinit
program neci_gpu_driver
    use iso_c_binding
    use neci_gpu_mod
    implicit none
    type(c_ptr) :: h
    type(ngpu_sys_t) :: sys
    type(ngpu_iter_in_t) :: inp
    type(ngpu_iter_out_t), target :: sm, last
    integer :: nsites, nel, niter, seed, truncinit, i, s, it, rc, nw
    real(c_double) :: u, tau, target_walkers, shift, hii, projE, growth, old_av, av
    real(c_double) :: sftdamp = 0.1_c_double, sum_shift
    integer :: stepssft = 10, n_sum
    logical :: single_part
    integer(c_int32_t), allocatable :: nbr(:), rh(:)
    integer(c_int64_t) :: ref(2), row(4)
    integer(c_int32_t) :: n_it
    integer(c_int64_t), parameter :: one = 1
    real(c_double) :: diag(1), offd(1), r
    character(len=32) :: arg

    nsites = 10; nel = 10; u = 4.0_c_double; tau = 0.01_c_double; niter = 2000
    target_walkers = 1.0e4_c_double; seed = 1234; truncinit = 0
    if (command_argument_count() >= 8) then
        call get_command_argument(1, arg); read(arg, *) nsites
        call get_command_argument(2, arg); read(arg, *) nel
        call get_command_argument(3, arg); read(arg, *) u
        call get_command_argument(4, arg); read(arg, *) tau
        call get_command_argument(5, arg); read(arg, *) niter
        call get_command_argument(6, arg); read(arg, *) target_walkers
        call get_command_argument(7, arg); read(arg, *) seed
        call get_command_argument(8, arg); read(arg, *) truncinit
    end if

    sys%nel = nel; sys%nbasis = 2 * nsites; sys%n_replicas = 1; sys%model = NGPU_MODEL_HUBBARD
    sys%t_trunc_initiator = truncinit; sys%t_all_real_coeff = 0; sys%t_real_spawn_cutoff = 1
    sys%flag_layout = 0; sys%initiator_walk_no = 3.0_c_double; sys%real_spawn_cutoff = 0.95_c_double
    sys%occupied_thresh = 1.0_c_double; sys%av_mc_excits = 1.0_c_double
    sys%max_walkers = int(10 * target_walkers, c_int64_t) + 1024
    sys%max_spawned = int(3 * target_walkers, c_int64_t) + 1024
    sys%n_hash_buckets = max(int(0.7_c_double * target_walkers, c_int64_t), one)
    sys%seed = seed; sys%pad = 0
    rc = ngpu_create(h, sys); call check(rc, "ngpu_create")

    ! periodic chain, neighbours (s-1, s+1)
    allocate(nbr(2 * nsites))
    do s = 1, nsites
        nbr(2 * s - 1) = merge(s - 1, nsites, s > 1)
        nbr(2 * s) = merge(s + 1, 1, s < nsites)
    end do
    rc = ngpu_set_lattice(h, nbr, nsites, 2, -1.0_c_double, u); call check(rc, "ngpu_set_lattice")

    ! RandomHash2: distinct integers in [1, 1000*nBasis] (fcimc_initialisation.fpp:890-915)
    allocate(rh(2 * nsites))
    call random_seed()
    i = 0
    do while (i < 2 * nsites)
        call random_number(r)
        s = int(2 * nsites * r * 1000) + 1
        if (i > 0) then
            if (any(rh(1:i) == s)) cycle
        end if
        i = i + 1
        rh(i) = s
    end do
    rc = ngpu_set_hash(h, rh); call check(rc, "ngpu_set_hash")

    ! Neel state: orbitals 1,4,5,8,9,...
    ref = 0
    do i = 1, nel
        s = 2 * i - 1
        if (mod(i, 2) == 0) s = s + 1
        ref((s - 1) / 64 + 1) = ibset(ref((s - 1) / 64 + 1), mod(s - 1, 64))
    end do
    rc = ngpu_set_reference(h, ref, 0.0_c_double); call check(rc, "ngpu_set_reference")
    rc = ngpu_diag_helement_batch(h, ref, one, diag, offd); call check(rc, "ngpu_diag_helement_batch")
    hii = diag(1)
    rc = ngpu_set_reference(h, ref, hii); call check(rc, "ngpu_set_reference")

    ! StartSinglePart: one walker on the reference, initiator flag (bit 13)
    row = 0
    row(1) = ref(1)
    row(2) = transfer(1.0_c_double, row(2))
    row(3) = ibset(0_c_int64_t, 13)
    rc = ngpu_upload_walkers(h, row, one); call check(rc, "ngpu_upload_walkers")

    shift = 0.0_c_double; single_part = .true.; old_av = 0.0_c_double
    sum_shift = 0.0_c_double; n_sum = 0
    write(*, '(a)') '#  iter        shift          proj.E       TotParts   dets'
    do it = 1, niter
        inp%tau = tau; inp%diag_sft = shift; inp%iter = it; inp%pad = 0
        rc = ngpu_iterate(h, inp, c_null_ptr); call check(rc, "ngpu_iterate")
        if (mod(it, stepssft) /= 0) cycle
        rc = ngpu_get_cycle_stats(h, sm, last, n_it); call check(rc, "ngpu_get_cycle_stats")
        av = sm%tot_parts(1) / n_it
        projE = 0.0_c_double
        if (abs(sm%hf_cyc(1)) > 1.0e-13_c_double) projE = sm%enum_cyc(1) / sm%hf_cyc(1)
        if (single_part) then
            if (last%tot_parts(1) > target_walkers) then
                single_part = .false.
                shift = projE
            end if
        else if (old_av > 0.0_c_double) then
            growth = av / old_av
            shift = shift - log(growth) * sftdamp / (tau * stepssft)
            if (it > niter / 3) then
                sum_shift = sum_shift + shift; n_sum = n_sum + 1
            end if
        end if
        old_av = av
        nw = int(last%tot_walkers)
        if (mod(it, 100) == 0) write(*, '(i7,2f16.8,f14.1,i8)') it, shift + hii, projE + hii, last%tot_parts(1), nw
    end do
    if (n_sum > 0) write(*, '(a,f16.8)') 'mean shift energy (last 2/3): ', sum_shift / n_sum + hii
    rc = ngpu_destroy(h)

contains

    subroutine check(code, where)
        integer(c_int), intent(in) :: code
        character(*), intent(in) :: where
        if (code /= 0) then
            write(*, '(a,a,i3)') where, ' failed with status ', code
            stop 1
        end if
    end subroutine
end program neci_gpu_driver
