! pdmd_f.f90 -- Fortran 2003 (ISO_C_BINDING) interface to the C ABI of the
! MI355X PRIME20 DMD engine (include/pdmd.h).  A Fortran driver keeps the
! reference's input files, parameter tables and output formats and calls the
! GPU event loop through these bindings.  Derived types mirror the C
! structs field for field (bind(C) gives them the C layout).
module pdmd_f
  use iso_c_binding
  implicit none

  type, bind(C) :: pdmd_system_desc
    integer(c_int32_t) :: nchains
    type(c_ptr) :: chain_len = c_null_ptr
    type(c_ptr) :: species = c_null_ptr
    type(c_ptr) :: restype = c_null_ptr
    type(c_ptr) :: hp = c_null_ptr
    real(c_double) :: boxl
    real(c_double) :: tstar
    integer(c_int64_t) :: seed
    integer(c_int32_t) :: canon
    integer(c_int32_t) :: list_capacity
    integer(c_int32_t) :: flags = 0
    integer(c_int32_t) :: pad = 0
    type(c_ptr) :: sv = c_null_ptr
    type(c_ptr) :: bptnr = c_null_ptr
  end type pdmd_system_desc

  type, bind(C) :: pdmd_stats_t
    integer(c_int64_t) :: coll, pair_events, numghosts, nupdates, nforced
    integer(c_int64_t) :: nevents(31)
    real(c_double) :: t, tfalse, interval, interval_max
    integer(c_int32_t) :: status, max_list, nbeads, pad
    real(c_double) :: avg_up, avg_dn
    integer(c_int64_t) :: phase_cycles(40)
    real(c_double) :: device_ms, rebuild_ms
    integer(c_int64_t) :: rebuilds
    integer(c_int32_t) :: waves, cell_factor, cells_axis, pad2
  end type pdmd_stats_t

  type, bind(C) :: pdmd_obs_t
    real(c_double) :: ered, tred, sumvel
    integer(c_int32_t) :: hb_alpha, hb_ii, hb_ij, pad
    real(c_double) :: ehh_ii, ehh_ij, rg_avg, e2e_avg
  end type pdmd_obs_t

  type, bind(C) :: pdmd_check_t
    integer(c_int64_t) :: overlaps, bonds_long, bonds_short
    integer(c_int32_t) :: first_i, first_j
    integer(c_int64_t) :: boundbad, unboundbad, ss_counted, ss_total
  end type pdmd_check_t

  type, bind(C) :: pdmd_event_t
    integer(c_int64_t) :: coll
    integer(c_int32_t) :: i, j, type, pad
    real(c_double) :: t
  end type pdmd_event_t

  interface
    type(c_ptr) function pdmd_version() bind(C, name='pdmd_version')
      import :: c_ptr
    end function
    type(c_ptr) function pdmd_last_error() bind(C, name='pdmd_last_error')
      import :: c_ptr
    end function
    integer(c_int) function pdmd_ctx_create(param_dir, ctx) bind(C, name='pdmd_ctx_create')
      import :: c_int, c_char, c_ptr
      character(kind=c_char), dimension(*), intent(in) :: param_dir
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function pdmd_ctx_destroy(ctx) bind(C, name='pdmd_ctx_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function pdmd_system_add(ctx, desc, sys_id) bind(C, name='pdmd_system_add')
      import :: c_int, c_ptr, c_int32_t, pdmd_system_desc
      type(c_ptr), value :: ctx
      type(pdmd_system_desc), intent(in) :: desc
      integer(c_int32_t), intent(out) :: sys_id
    end function
    integer(c_int) function pdmd_system_remove(ctx, sys_id) bind(C, name='pdmd_system_remove')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: sys_id
    end function
    integer(c_int64_t) function pdmd_system_bytes(desc) bind(C, name='pdmd_system_bytes')
      import :: c_int64_t, pdmd_system_desc
      type(pdmd_system_desc), intent(in) :: desc
    end function
    integer(c_int64_t) function pdmd_system_bytes_ctx(ctx, desc) bind(C, name='pdmd_system_bytes_ctx')
      import :: c_int64_t, c_ptr, pdmd_system_desc
      type(c_ptr), value :: ctx
      type(pdmd_system_desc), intent(in) :: desc
    end function
    integer(c_int) function pdmd_advance(ctx, nsys, ids, nevents) bind(C, name='pdmd_advance')
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsys
      integer(c_int32_t), intent(in) :: ids(*)
      integer(c_int64_t), intent(in) :: nevents(*)
    end function
    integer(c_int) function pdmd_set_waves(ctx, waves_per_system) bind(C, name='pdmd_set_waves')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: waves_per_system
    end function
    integer(c_int) function pdmd_rebuild(ctx, nsys, ids) bind(C, name='pdmd_rebuild')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsys
      integer(c_int32_t), intent(in) :: ids(*)
    end function
    integer(c_int) function pdmd_set_graph_launch(ctx, on) bind(C, name='pdmd_set_graph_launch')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: on
    end function
    integer(c_int) function pdmd_advance_to_output(ctx, nsys, ids, max_nevents) bind(C, name='pdmd_advance_to_output')
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsys
      integer(c_int32_t), intent(in) :: ids(*)
      integer(c_int64_t), intent(in) :: max_nevents(*)
    end function
    ! array arguments as pointers so that any of them may be c_null_ptr
    integer(c_int) function pdmd_download(ctx, sys_id, sv, identity, bptnr, extra_repuls, tim, nptnr, coltype) &
        bind(C, name='pdmd_download')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: sys_id
      type(c_ptr), value :: sv, identity, bptnr, extra_repuls, tim, nptnr, coltype
    end function
    integer(c_int) function pdmd_check(ctx, sys_id, sv, out) bind(C, name='pdmd_check')
      import :: c_int, c_ptr, c_int32_t, pdmd_check_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: sys_id
      type(c_ptr), value :: sv
      type(pdmd_check_t), intent(out) :: out
    end function
    integer(c_int) function pdmd_stats(ctx, sys_id, out) bind(C, name='pdmd_stats')
      import :: c_int, c_ptr, c_int32_t, pdmd_stats_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: sys_id
      type(pdmd_stats_t), intent(out) :: out
    end function
    integer(c_int) function pdmd_observables(ctx, sys_id, out) bind(C, name='pdmd_observables')
      import :: c_int, c_ptr, c_int32_t, pdmd_obs_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: sys_id
      type(pdmd_obs_t), intent(out) :: out
    end function
    integer(c_int) function pdmd_generate(param_dir, nchains, chain_len, density, tstar, seed, restype, hp, sv, boxl) &
        bind(C, name='pdmd_generate')
      import :: c_int, c_char, c_int32_t, c_double, c_int64_t
      character(kind=c_char), dimension(*), intent(in) :: param_dir
      integer(c_int32_t), value :: nchains
      integer(c_int32_t), intent(in) :: chain_len(*)
      real(c_double), value :: density, tstar
      integer(c_int64_t), value :: seed
      integer(c_int32_t), intent(out) :: restype(*), hp(*)
      real(c_double), intent(out) :: sv(*)
      real(c_double), intent(out) :: boxl
    end function
    integer(c_int) function pdmd_generate_seq(param_dir, nchains, chain_len, restype_in, density, tstar, seed, &
        restype, hp, sv, boxl) bind(C, name='pdmd_generate_seq')
      import :: c_int, c_char, c_int32_t, c_double, c_int64_t
      character(kind=c_char), dimension(*), intent(in) :: param_dir
      integer(c_int32_t), value :: nchains
      integer(c_int32_t), intent(in) :: chain_len(*), restype_in(*)
      real(c_double), value :: density, tstar
      integer(c_int64_t), value :: seed
      integer(c_int32_t), intent(out) :: restype(*), hp(*)
      real(c_double), intent(out) :: sv(*)
      real(c_double), intent(out) :: boxl
    end function
    integer(c_int) function pdmd_pair_codes(desc, n, a, b, code) bind(C, name='pdmd_pair_codes')
      import :: c_int, c_int32_t, pdmd_system_desc
      type(pdmd_system_desc), intent(in) :: desc
      integer(c_int32_t), value :: n
      integer(c_int32_t), intent(in) :: a(*), b(*)
      integer(c_int32_t), intent(out) :: code(*)
    end function
    function c_strlen(s) bind(C, name='strlen')
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
      integer(c_size_t) :: c_strlen
    end function
  end interface

contains

  ! Fortran copy of a NUL-terminated C string
  function f_string(p) result(s)
    type(c_ptr), intent(in) :: p
    character(len=:), allocatable :: s
    character(kind=c_char), pointer :: a(:)
    integer :: n, k
    if (.not. c_associated(p)) then
      s = ''
      return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, a, [n])
    allocate(character(len=n) :: s)
    do k = 1, n
      s(k:k) = a(k)
    end do
  end function f_string

  ! the engine's last error text, for messages before exit(-1) as the
  ! reference prints them
  function pdmd_errmsg() result(s)
    character(len=:), allocatable :: s
    s = f_string(pdmd_last_error())
  end function pdmd_errmsg

end module pdmd_f
