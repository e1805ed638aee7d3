! neci_gpu_mod -- ISO_C_BINDING interface to the GPU iteration engine
! (include/neci_gpu.h).  This is the module a NECI host build uses in place of
! the body of PerformFCIMCycPar (src/FciMCPar.F90:1226-1873): integrals and
! keywords stay on the host, one ngpu_iterate call per iteration, the update
! cycle (update_shift) stays in Fortran.
module neci_gpu_mod
    use iso_c_binding
    implicit none
    private

    integer(c_int), parameter, public :: NGPU_MAX_REPLICAS = 16
    integer(c_int), parameter, public :: NGPU_MODEL_HUBBARD = 0, NGPU_MODEL_MOLECULAR = 1
    integer(c_int), parameter, public :: NGPU_SINGLES_HEATBATH = 1
    integer(c_int), parameter, public :: NGPU_PART_UNIF_UNIF = 0, NGPU_PART_UNIF_FAST = 1
    integer(c_int), parameter, public :: NGPU_PART_UNIF_FULL = 2, NGPU_PART_FULL_FULL = 3
    integer(c_int), parameter, public :: NGPU_HOLES_FAST_FAST = 0, NGPU_HOLES_FULL_FULL = 1
    integer(c_int), parameter, public :: NGPU_CORE_BUILD_AUTO = 0, NGPU_CORE_BUILD_JOIN = 1, NGPU_CORE_BUILD_SCAN = 2
    integer(c_int), parameter, public :: NGPU_GROUP_CORE = 1, NGPU_GROUP_SHARED_CORE = 2
    integer(c_int), parameter, public :: NGPU_CHECK_TABLE = 1, NGPU_CHECK_PGEN = 2, NGPU_CHECK_SELFTEST = 4

    type, bind(c), public :: ngpu_sys_t
        integer(c_int32_t) :: nel, nbasis, n_replicas, model
        integer(c_int32_t) :: t_trunc_initiator, t_all_real_coeff, t_real_spawn_cutoff, flag_layout
        real(c_double) :: initiator_walk_no, real_spawn_cutoff, occupied_thresh, av_mc_excits
        integer(c_int64_t) :: max_walkers, max_spawned, n_hash_buckets
        integer(c_int32_t) :: seed, pad
    end type

    type, bind(c), public :: ngpu_iter_in_t
        real(c_double) :: tau
        real(c_double) :: diag_sft(NGPU_MAX_REPLICAS)
        integer(c_int32_t) :: iter, pad
    end type

    type, bind(c), public :: ngpu_iter_out_t
        real(c_double) :: nborn(NGPU_MAX_REPLICAS), ndied(NGPU_MAX_REPLICAS), nannihil(NGPU_MAX_REPLICAS)
        real(c_double) :: naborted(NGPU_MAX_REPLICAS), nremoved(NGPU_MAX_REPLICAS)
        real(c_double) :: tot_parts(NGPU_MAX_REPLICAS), norm_sq(NGPU_MAX_REPLICAS)
        real(c_double) :: enum_cyc(NGPU_MAX_REPLICAS), hf_cyc(NGPU_MAX_REPLICAS), noat_doubs(NGPU_MAX_REPLICAS)
        real(c_double) :: spawn_from_sing(NGPU_MAX_REPLICAS), inst_no_at_hf(NGPU_MAX_REPLICAS)
        real(c_double) :: no_init_dets(NGPU_MAX_REPLICAS), no_init_walk(NGPU_MAX_REPLICAS)
        real(c_double) :: no_noninit_dets(NGPU_MAX_REPLICAS), no_noninit_walk(NGPU_MAX_REPLICAS)
        real(c_double) :: n_valid_excits, n_invalid_excits, max_cyc_spawn
        real(c_double) :: bloom_count(3), bloom_size(3)
        integer(c_int64_t) :: tot_walkers, n_attempts, n_spawned, n_unique, n_new
        integer(c_int32_t) :: err, pad
    end type

    public :: ngpu_create, ngpu_destroy, ngpu_set_integrals, ngpu_set_lattice, ngpu_set_excitgen
    public :: ngpu_set_hash, ngpu_set_reference, ngpu_upload_walkers, ngpu_download_walkers
    public :: ngpu_iterate, ngpu_get_stats, ngpu_get_cycle_stats, ngpu_find_walker_hash_batch
    public :: ngpu_diag_helement_batch, ngpu_set_excitgen_uniform, ngpu_set_core
    public :: ngpu_set_graph, ngpu_batch_create, ngpu_batch_iterate, ngpu_batch_destroy
    public :: ngpu_fcidump_header, ngpu_read_fcidump, ngpu_set_spawn_precision, ngpu_get_hash_stats
    public :: ngpu_get_error, ngpu_set_excitgen_pchb, ngpu_download_walkers_dev, ngpu_most_populated
    public :: ngpu_select_core, ngpu_set_core_build, ngpu_get_core_hamil, ngpu_batch_get_groups
    public :: ngpu_batch_set_graph, ngpu_batch_set_profiling, ngpu_batch_get_group_flags
    public :: ngpu_set_checks, ngpu_get_check_counts

    interface
        integer(c_int) function ngpu_create(h, sys) bind(c, name="ngpu_create")
            import :: c_int, c_ptr, ngpu_sys_t
            type(c_ptr), intent(out) :: h
            type(ngpu_sys_t), intent(in) :: sys
        end function
        integer(c_int) function ngpu_destroy(h) bind(c, name="ngpu_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
        end function
        integer(c_int) function ngpu_set_integrals(h, tmat2d, umat, umat_len, ecore) &
            bind(c, name="ngpu_set_integrals")
            import :: c_int, c_ptr, c_double, c_int64_t
            type(c_ptr), value :: h
            real(c_double), intent(in) :: tmat2d(*), umat(*)
            integer(c_int64_t), value :: umat_len
            real(c_double), value :: ecore
        end function
        integer(c_int) function ngpu_set_lattice(h, nbr, nsites, maxnbr, bhub, uhub) &
            bind(c, name="ngpu_set_lattice")
            import :: c_int, c_ptr, c_double, c_int32_t
            type(c_ptr), value :: h
            integer(c_int32_t), intent(in) :: nbr(*)
            integer(c_int32_t), value :: nsites, maxnbr
            real(c_double), value :: bhub, uhub
        end function
        integer(c_int) function ngpu_set_excitgen(h, singles, particles, psingles, pparallel) &
            bind(c, name="ngpu_set_excitgen")
            import :: c_int, c_ptr, c_double, c_int32_t
            type(c_ptr), value :: h
            integer(c_int32_t), value :: singles, particles
            real(c_double), value :: psingles, pparallel
        end function
        integer(c_int) function ngpu_set_excitgen_uniform(h, psingles) &
            bind(c, name="ngpu_set_excitgen_uniform")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), value :: psingles
        end function
        integer(c_int) function ngpu_set_spawn_precision(h, fp32) bind(c, name="ngpu_set_spawn_precision")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: h
            integer(c_int32_t), value :: fp32
        end function
        integer(c_int) function ngpu_get_error(h, status, iter) bind(c, name="ngpu_get_error")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: h
            integer(c_int32_t), intent(out) :: status, iter
        end function
        integer(c_int) function ngpu_get_hash_stats(h, live, tombstones, entries, rebuilds) &
            bind(c, name="ngpu_get_hash_stats")
            import :: c_int, c_ptr, c_int64_t
            type(c_ptr), value :: h
            integer(c_int64_t), intent(out) :: live, tombstones, entries, rebuilds
        end function
        integer(c_int) function ngpu_set_core(h, dets, ncore, fp32_values, nnz) bind(c, name="ngpu_set_core")
            import :: c_int, c_ptr, c_int64_t, c_int32_t
            type(c_ptr), value :: h
            integer(c_int64_t), intent(in) :: dets(*)
            integer(c_int64_t), value :: ncore
            integer(c_int32_t), value :: fp32_values
            integer(c_int64_t), intent(out) :: nnz
        end function
        integer(c_int) function ngpu_set_graph(h, on) bind(c, name="ngpu_set_graph")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: h
            integer(c_int32_t), value :: on
        end function
        integer(c_int) function ngpu_batch_create(b, hs, n) bind(c, name="ngpu_batch_create")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), intent(out) :: b
            type(c_ptr), intent(in) :: hs(*)
            integer(c_int32_t), value :: n
        end function
        integer(c_int) function ngpu_batch_iterate(b, ins) bind(c, name="ngpu_batch_iterate")
            import :: c_int, c_ptr, ngpu_iter_in_t
            type(c_ptr), value :: b
            type(ngpu_iter_in_t), intent(in) :: ins(*)
        end function
        integer(c_int) function ngpu_batch_destroy(b) bind(c, name="ngpu_batch_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: b
        end function
        integer(c_int) function ngpu_fcidump_header(path, norb, nelec, ms2, isym, orbsym, orbsym_cap) &
            bind(c, name="ngpu_fcidump_header")
            import :: c_int, c_char, c_int32_t
            character(kind=c_char), intent(in) :: path(*)
            integer(c_int32_t), intent(out) :: norb, nelec, ms2, isym
            integer(c_int32_t), intent(out) :: orbsym(*)
            integer(c_int32_t), value :: orbsym_cap
        end function
        integer(c_int) function ngpu_read_fcidump(path, umat_eps, tmat2d, nbasis, umat, umat_len, ecore, n_two) &
            bind(c, name="ngpu_read_fcidump")
            import :: c_int, c_char, c_double, c_int64_t
            character(kind=c_char), intent(in) :: path(*)
            real(c_double), value :: umat_eps
            real(c_double), intent(out) :: tmat2d(*)
            integer(c_int64_t), value :: nbasis
            real(c_double), intent(out) :: umat(*)
            integer(c_int64_t), value :: umat_len
            real(c_double), intent(out) :: ecore
            integer(c_int64_t), intent(out) :: n_two
        end function
        integer(c_int) function ngpu_set_hash(h, random_hash2) bind(c, name="ngpu_set_hash")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: h
            integer(c_int32_t), intent(in) :: random_hash2(*)
        end function
        integer(c_int) function ngpu_set_reference(h, ilut_ref, hii) bind(c, name="ngpu_set_reference")
            import :: c_int, c_ptr, c_int64_t, c_double
            type(c_ptr), value :: h
            integer(c_int64_t), intent(in) :: ilut_ref(*)
            real(c_double), value :: hii
        end function
        integer(c_int) function ngpu_upload_walkers(h, rows, n) bind(c, name="ngpu_upload_walkers")
            import :: c_int, c_ptr, c_int64_t
            type(c_ptr), value :: h
            integer(c_int64_t), intent(in) :: rows(*)
            integer(c_int64_t), value :: n
        end function
        integer(c_int) function ngpu_download_walkers(h, rows, diag_h, offdiag_h, cap, n) &
            bind(c, name="ngpu_download_walkers")
            import :: c_int, c_ptr, c_int64_t, c_double
            type(c_ptr), value :: h
            integer(c_int64_t), intent(out) :: rows(*)
            real(c_double), intent(out) :: diag_h(*), offdiag_h(*)
            integer(c_int64_t), value :: cap
            integer(c_int64_t), intent(out) :: n
        end function
        integer(c_int) function ngpu_iterate(h, inp, out) bind(c, name="ngpu_iterate")
            import :: c_int, c_ptr, ngpu_iter_in_t
            type(c_ptr), value :: h
            type(ngpu_iter_in_t), intent(in) :: inp
            type(c_ptr), value :: out
        end function
        integer(c_int) function ngpu_get_stats(h, out) bind(c, name="ngpu_get_stats")
            import :: c_int, c_ptr, ngpu_iter_out_t
            type(c_ptr), value :: h
            type(ngpu_iter_out_t), intent(out) :: out
        end function
        integer(c_int) function ngpu_get_cycle_stats(h, sum, last, n_iters) bind(c, name="ngpu_get_cycle_stats")
            import :: c_int, c_ptr, ngpu_iter_out_t, c_int32_t
            type(c_ptr), value :: h
            type(ngpu_iter_out_t), intent(out) :: sum, last
            integer(c_int32_t), intent(out) :: n_iters
        end function
        integer(c_int) function ngpu_find_walker_hash_batch(h, dets, n, buckets) &
            bind(c, name="ngpu_find_walker_hash_batch")
            import :: c_int, c_ptr, c_int64_t, c_int32_t
            type(c_ptr), value :: h
            integer(c_int64_t), intent(in) :: dets(*)
            integer(c_int64_t), value :: n
            integer(c_int32_t), intent(out) :: buckets(*)
        end function
        integer(c_int) function ngpu_diag_helement_batch(h, dets, n, diag, offdiag) &
            bind(c, name="ngpu_diag_helement_batch")
            import :: c_int, c_ptr, c_int64_t, c_double
            type(c_ptr), value :: h
            integer(c_int64_t), intent(in) :: dets(*)
            integer(c_int64_t), value :: n
            real(c_double), intent(out) :: diag(*), offdiag(*)
        end function
        integer(c_int) function ngpu_set_excitgen_pchb(h, singles, particles, holes, psingles, pparallel) &
            bind(c, name="ngpu_set_excitgen_pchb")
            import :: c_int, c_ptr, c_double, c_int32_t
            type(c_ptr), value :: h
            integer(c_int32_t), value :: singles, particles, holes
            real(c_double), value :: psingles, pparallel
        end function
        ! rows, diag_h, offdiag_h and n are device addresses (diag_h, offdiag_h may be c_null_ptr)
        integer(c_int) function ngpu_download_walkers_dev(h, rows, diag_h, offdiag_h, cap, n) &
            bind(c, name="ngpu_download_walkers_dev")
            import :: c_int, c_ptr, c_int64_t
            type(c_ptr), value :: h, rows, diag_h, offdiag_h
            integer(c_int64_t), value :: cap
            type(c_ptr), value :: n
        end function
        integer(c_int) function ngpu_most_populated(h, target, dets, n) bind(c, name="ngpu_most_populated")
            import :: c_int, c_ptr, c_int64_t
            type(c_ptr), value :: h
            integer(c_int64_t), value :: target
            integer(c_int64_t), intent(out) :: dets(*)
            integer(c_int64_t), intent(out) :: n
        end function
        integer(c_int) function ngpu_select_core(h, target, fp32_values, dets, ncore, nnz) &
            bind(c, name="ngpu_select_core")
            import :: c_int, c_ptr, c_int64_t, c_int32_t
            type(c_ptr), value :: h
            integer(c_int64_t), value :: target
            integer(c_int32_t), value :: fp32_values
            integer(c_int64_t), intent(out) :: dets(*)
            integer(c_int64_t), intent(out) :: ncore, nnz
        end function
        integer(c_int) function ngpu_set_core_build(h, method) bind(c, name="ngpu_set_core_build")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: h
            integer(c_int32_t), value :: method
        end function
        integer(c_int) function ngpu_get_core_hamil(h, rowptr, cols, vals, cap, nnz) &
            bind(c, name="ngpu_get_core_hamil")
            import :: c_int, c_ptr, c_int64_t, c_int32_t, c_double
            type(c_ptr), value :: h
            integer(c_int64_t), intent(out) :: rowptr(*)
            integer(c_int32_t), intent(out) :: cols(*)
            real(c_double), intent(out) :: vals(*)
            integer(c_int64_t), value :: cap
            integer(c_int64_t), intent(out) :: nnz
        end function
        integer(c_int) function ngpu_batch_get_groups(b, group_of_engine, n_groups) &
            bind(c, name="ngpu_batch_get_groups")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: b
            integer(c_int32_t), intent(out) :: group_of_engine(*)
            integer(c_int32_t), intent(out) :: n_groups
        end function
        integer(c_int) function ngpu_batch_get_group_flags(b, group, flags) bind(c, name="ngpu_batch_get_group_flags")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: b
            integer(c_int32_t), value :: group
            integer(c_int32_t), intent(out) :: flags
        end function
        integer(c_int) function ngpu_set_checks(h, mask) bind(c, name="ngpu_set_checks")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: h
            integer(c_int32_t), value :: mask
        end function
        integer(c_int) function ngpu_get_check_counts(h, counts, n) bind(c, name="ngpu_get_check_counts")
            import :: c_int, c_ptr, c_int64_t, c_int32_t
            type(c_ptr), value :: h
            integer(c_int64_t), intent(out) :: counts(*)
            integer(c_int32_t), value :: n
        end function
        integer(c_int) function ngpu_batch_set_graph(b, on) bind(c, name="ngpu_batch_set_graph")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: b
            integer(c_int32_t), value :: on
        end function
        integer(c_int) function ngpu_batch_set_profiling(b, on) bind(c, name="ngpu_batch_set_profiling")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: b
            integer(c_int32_t), value :: on
        end function
    end interface
end module neci_gpu_mod
