! pdmd_driver.f90 -- Fortran host driver of the MI355X PRIME20 DMD engine.
!
! Keeps the reference driver's interface (parallel-dmd-PRIME20/code/
! main.F90): stdin is a temp_XXX file (line 1 T*, line 2 the number of
! calendar entries to commit), inputs are read from parameters/ and
! results/ in the reference's formats, and the outputs are the reference's
! results/runNNNN.energy (format 22223 of main.F90:1380 at every output
! pseudo-event), runNNNN.config and runNNNN.bptnr (appended unformatted
! records, config.f:22-39) and runNNNN.lastvel.  The event loop itself runs
! on the GPU through the C ABI (module pdmd_f).
!
! What the reference fixes at compile time (-Dnop1 -Dnop2 -Dchnln1 ...
! -Dcanon -Drunr/-Drunh) is read from the namelist file pdmd.nml:
!
!   &pdmd_run
!     mode = 'run'          ! 'run', or 'generate' to write a seeded start
!     nchains1 = 8, chnln1 = 16, nchains2 = 0, chnln2 = 0
!     start = 'helix'       ! 'helix': parameters/helix.inp (-Drunh)
!                           ! 'restart': results/run<prev>.config/.lastvel/.bptnr (-Drunr)
!     boxl = 0.0            ! box length in Angstrom for 'restart'
!     canon = 1             ! Andersen thermostat (-Dcanon)
!     seed = 1              ! project RNG seed (replaces IFPORT srand/drandm)
!     param_dir = 'parameters'
!     density = 0.01, gen_seed = 1   ! 'generate' only
!     waves = 1             ! wavefronts per system (1 to 8; same results)
!     no_hbs = 0            ! 1: build flag -Dno_hbs (no N-C H-bond wells)
!     no_h = 0              ! 1: build flag -Dno_h (no side-chain wells between chains)
!     chaptype = 1          ! 2: build flag -Dchaptype=2 (second species unbonded)
!     debugging = 0         ! 1: -Ddebugging checks at every output event
!   /
!
! 'generate' writes parameters/identity.inp, hp1.inp, hp2.inp,
! firstside1.data, firstside2.data and helix.inp (inputinfo.f:28-66,
! 105-262 formats) for a seeded random system at the namelist's sizes, T*
! and density, so that both this driver and the reference can start from it.
module pdmd_driver_m
  use iso_c_binding
  use pdmd_f
  implicit none
  character(len=16) :: mode = 'run', start = 'helix'
  character(len=256) :: param_dir = 'parameters'
  integer :: nchains1 = 8, chnln1 = 16, nchains2 = 0, chnln2 = 0, canon = 1, list_capacity = 0, waves = 1
  integer :: no_hbs = 0, no_h = 0, debugging = 0, chaptype = 1
  integer(c_int64_t) :: seed = 1, gen_seed = 1
  real(c_double) :: boxl = 0.0d0, density = 0.01d0
  namelist /pdmd_run/ mode, nchains1, chnln1, nchains2, chnln2, start, boxl, canon, seed, param_dir, density, &
                      gen_seed, list_capacity, waves, no_hbs, no_h, debugging, chaptype

  real(c_double) :: tstar, setemp, sigma1
  integer(c_int64_t) :: ncoll
  integer :: u, ios

contains

  integer function numbeads(L)
    integer, intent(in) :: L
    numbeads = 4 * L
  end function numbeads

  subroutine die(msg)
    character(len=*), intent(in) :: msg
    write(6, *) trim(msg)
    call exit(-1)
  end subroutine die

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= 0) call die(what // ': ' // pdmd_errmsg())
  end subroutine check

  ! chains in reference order: species 1 then species 2
  subroutine chain_table(clen, spec)
    integer(c_int32_t), allocatable, intent(out) :: clen(:), spec(:)
    integer :: k
    allocate(clen(nchains1 + nchains2), spec(nchains1 + nchains2))
    do k = 1, nchains1
      clen(k) = chnln1
      spec(k) = 0
    end do
    do k = 1, nchains2
      clen(nchains1 + k) = chnln2
      spec(nchains1 + k) = 1
    end do
  end subroutine chain_table

  ! ---- 'generate': a seeded start in the reference's input formats
  subroutine generate_start()
    integer(c_int32_t), allocatable :: clen(:), spec(:), rt(:), hp(:), seq(:), rt1(:), hp1(:)
    real(c_double), allocatable :: sv(:), sv1(:)
    real(c_double) :: b, b1
    integer :: nch, nres, N, k, c, r, off
    character(len=:), allocatable :: pd
    pd = trim(param_dir) // c_null_char
    call chain_table(clen, spec)
    nch = nchains1 + nchains2
    nres = sum(clen)
    N = 4 * nres
    ! one sequence per species, drawn from the seed, shared by its chains
    allocate(seq(nres))
    do k = 1, 2
      if ((k == 1 .and. nchains1 == 0) .or. (k == 2 .and. nchains2 == 0)) cycle
      if (k == 1) then
        allocate(rt1(chnln1), hp1(chnln1), sv1(24 * chnln1))
        call check(pdmd_generate(pd, 1_c_int32_t, [int(chnln1, c_int32_t)], density, tstar, gen_seed + k, &
                                 rt1, hp1, sv1, b1), 'pdmd_generate')
        do c = 1, nchains1
          seq((c - 1) * chnln1 + 1:c * chnln1) = rt1
        end do
        deallocate(rt1, hp1, sv1)
      else
        allocate(rt1(chnln2), hp1(chnln2), sv1(24 * chnln2))
        call check(pdmd_generate(pd, 1_c_int32_t, [int(chnln2, c_int32_t)], density, tstar, gen_seed + k, &
                                 rt1, hp1, sv1, b1), 'pdmd_generate')
        off = nchains1 * chnln1
        do c = 1, nchains2
          seq(off + (c - 1) * chnln2 + 1:off + c * chnln2) = rt1
        end do
        deallocate(rt1, hp1, sv1)
      end if
    end do
    allocate(rt(nres), hp(nres), sv(6 * N))
    call check(pdmd_generate_seq(pd, int(nch, c_int32_t), clen, seq, density, tstar, gen_seed, rt, hp, sv, b), &
               'pdmd_generate_seq')
    ! identity.inp: one chain of each species, bead order Calpha, N, C, R
    open(newunit=u, file=trim(param_dir) // '/identity.inp', status='replace')
    if (nchains1 > 0) write(u, '(*(i4))') (2, k = 1, chnln1), (1, k = 1, chnln1), (4, k = 1, chnln1), &
                                          (rt(r), r = 1, chnln1)
    if (nchains2 > 0) write(u, '(*(i4))') (2, k = 1, chnln2), (1, k = 1, chnln2), (4, k = 1, chnln2), &
                                          (rt(nchains1 * chnln1 + r), r = 1, chnln2)
    close(u)
    call write_hp('/hp1.inp', chnln1, 0)
    call write_hp('/hp2.inp', max(chnln2, 1), nchains1 * chnln1)
    open(newunit=u, file=trim(param_dir) // '/firstside1.data', status='replace')
    write(u, '(*(i3))') (1, k = 1, chnln1)
    close(u)
    open(newunit=u, file=trim(param_dir) // '/firstside2.data', status='replace')
    write(u, '(*(i3))') (1, k = 1, max(chnln2, 1))
    close(u)
    ! helix.inp (-Drunh): bead count, box, positions in Angstrom, velocities
    open(newunit=u, file=trim(param_dir) // '/helix.inp', status='replace')
    write(u, *) N
    write(u, '(es25.17)') b
    write(u, '(es25.17)') (sv(6 * (k - 1) + 1) * b, k = 1, N)
    write(u, '(es25.17)') (sv(6 * (k - 1) + 2) * b, k = 1, N)
    write(u, '(es25.17)') (sv(6 * (k - 1) + 3) * b, k = 1, N)
    write(u, '(es25.17)') (sv(6 * (k - 1) + 4), k = 1, N)
    write(u, '(es25.17)') (sv(6 * (k - 1) + 5), k = 1, N)
    write(u, '(es25.17)') (sv(6 * (k - 1) + 6), k = 1, N)
    close(u)
    write(6, '(a,i8,a,f12.4,a)') 'generated ', N, ' beads, box ', b, ' Angstrom'
  contains
    subroutine write_hp(name, L, roff)
      character(len=*), intent(in) :: name
      integer, intent(in) :: L, roff
      integer :: q
      open(newunit=u, file=trim(param_dir) // name, status='replace')
      write(u, '(*(i3))') (0, q = 1, 3 * L), (hp(min(roff + q, nres)), q = 1, L)
      close(u)
    end subroutine write_hp
  end subroutine generate_start

  ! ---- 'run': the event loop on the GPU with the reference's outputs
  subroutine run()
    integer(c_int32_t), allocatable, target :: clen(:), spec(:), rt(:), hp(:), bptnr(:), ident(:)
    integer, allocatable :: aa(:), hpb1(:), hpb2(:)
    real(c_double), allocatable, target :: sv(:)
    real(c_double), allocatable :: x(:), y(:), z(:)
    type(pdmd_system_desc) :: desc
    type(pdmd_stats_t) :: st
    type(pdmd_obs_t) :: ob
    type(pdmd_check_t) :: chk
    integer(c_int64_t) :: boundbad, unboundbad
    type(c_ptr) :: ctx
    integer(c_int32_t) :: sid
    integer(c_int64_t) :: coll, coll_in, before, left
    real(c_double) :: t_in, tsim
    character(len=4) :: digits, digits_pre
    character(len=:), allocatable :: pd
    integer :: nch, nres, N, nb1, nb2, k, c, r, off, ue
    logical :: ex

    call chain_table(clen, spec)
    nch = nchains1 + nchains2
    nres = sum(clen)
    N = 4 * nres
    nb1 = numbeads(chnln1)
    nb2 = numbeads(max(chnln2, 1))
    ! files_opn.f:17-41: first free results/runNNNN.energy, previous run
    k = 0
    do
      write(digits, '(i4.4)') k
      inquire(file='results/run' // digits // '.energy', exist=ex)
      if (.not. ex) exit
      k = k + 1
      if (k > 9999) call die('file name error. too many output files.')
    end do
    write(digits_pre, '(i4.4)') max(k - 1, 0)
    ! protein.data line 1: nitrogen bead diameter (time unit of the output)
    open(newunit=u, file=trim(param_dir) // '/protein.data', status='old', iostat=ios)
    if (ios /= 0) call die('cannot open ' // trim(param_dir) // '/protein.data')
    read(u, *) sigma1
    close(u)
    ! identity.inp / hp1.inp / hp2.inp (inputinfo.f:226-262)
    allocate(aa(nb1 + merge(nb2, 0, nchains2 > 0)), hpb1(nb1), hpb2(nb2))
    open(newunit=u, file=trim(param_dir) // '/identity.inp', status='old', iostat=ios)
    if (ios /= 0) call die('cannot open identity.inp')
    read(u, *) aa
    close(u)
    open(newunit=u, file=trim(param_dir) // '/hp1.inp', status='old', iostat=ios)
    if (ios /= 0) call die('cannot open hp1.inp')
    read(u, *) hpb1
    close(u)
    hpb2 = 1
    if (nchains2 > 0) then
      open(newunit=u, file=trim(param_dir) // '/hp2.inp', status='old', iostat=ios)
      if (ios /= 0) call die('cannot open hp2.inp')
      read(u, *) hpb2
      close(u)
    end if
    allocate(rt(nres), hp(nres))
    do c = 1, nchains1
      do r = 1, chnln1
        rt((c - 1) * chnln1 + r) = aa(3 * chnln1 + r)
        hp((c - 1) * chnln1 + r) = hpb1(3 * chnln1 + r)
      end do
    end do
    off = nchains1 * chnln1
    do c = 1, nchains2
      do r = 1, chnln2
        rt(off + (c - 1) * chnln2 + r) = aa(nb1 + 3 * chnln2 + r)
        hp(off + (c - 1) * chnln2 + r) = hpb2(3 * chnln2 + r)
      end do
    end do
    ! starting state
    allocate(sv(6 * N), x(N), y(N), z(N), bptnr(N), ident(N))
    bptnr = 0
    coll_in = 0
    if (trim(start) == 'helix') then
      ! inputinfo.f:30-66 with scale_down.f:18-27
      open(newunit=u, file=trim(param_dir) // '/helix.inp', status='old', iostat=ios)
      if (ios /= 0) call die('cannot open helix.inp')
      read(u, *) k
      if (k /= N) call die('nop error in reading starting pos and vel')
      read(u, *) boxl
      read(u, *) x, y, z
      do k = 1, N
        sv(6 * (k - 1) + 1) = x(k) / boxl
        sv(6 * (k - 1) + 2) = y(k) / boxl
        sv(6 * (k - 1) + 3) = z(k) / boxl
      end do
      read(u, *) x, y, z
      do k = 1, N
        sv(6 * (k - 1) + 4) = x(k)
        sv(6 * (k - 1) + 5) = y(k)
        sv(6 * (k - 1) + 6) = z(k)
      end do
      close(u)
    else
      ! -Drunr: last records of the previous run (inputinfo.f:69-101,
      ! main.F90:240-260)
      if (.not. (boxl > 0.d0)) call die('restart start needs boxl in pdmd.nml')
      open(newunit=u, file='results/run' // digits_pre // '.config', status='old', form='unformatted', iostat=ios)
      if (ios /= 0) call die('cannot open results/run' // digits_pre // '.config')
      do
        read(u, iostat=ios) coll_in, t_in, x, y, z
        if (ios /= 0) exit
        do k = 1, N
          sv(6 * (k - 1) + 1) = x(k) - dnint(x(k))
          sv(6 * (k - 1) + 2) = y(k) - dnint(y(k))
          sv(6 * (k - 1) + 3) = z(k) - dnint(z(k))
        end do
      end do
      close(u)
      open(newunit=u, file='results/run' // digits_pre // '.lastvel', status='old', form='unformatted', iostat=ios)
      if (ios /= 0) call die('cannot open results/run' // digits_pre // '.lastvel')
      read(u) coll_in, x, y, z
      close(u)
      do k = 1, N
        sv(6 * (k - 1) + 4) = x(k)
        sv(6 * (k - 1) + 5) = y(k)
        sv(6 * (k - 1) + 6) = z(k)
      end do
      open(newunit=u, file='results/run' // digits_pre // '.bptnr', status='old', form='unformatted', iostat=ios)
      if (ios == 0) then
        do
          read(u, iostat=ios) coll_in, ident
          if (ios /= 0) exit
          bptnr = ident
        end do
        close(u)
      end if
    end if

    if (trim(mode) == 'codes') then
      ! host-only check of the inputs as the engine sees them: pair codes
      ! (ev_code of make_code.f) of every pair of the first chain
      call codes_out()
      return
    end if
    desc%nchains = int(nch, c_int32_t)
    desc%chain_len = c_loc(clen)
    desc%species = c_loc(spec)
    desc%restype = c_loc(rt)
    desc%hp = c_loc(hp)
    desc%boxl = boxl
    desc%tstar = tstar
    desc%seed = seed
    desc%canon = int(canon, c_int32_t)
    desc%list_capacity = int(list_capacity, c_int32_t)
    desc%flags = merge(1_c_int32_t, 0_c_int32_t, no_hbs /= 0) + merge(2_c_int32_t, 0_c_int32_t, no_h /= 0) &
                 + merge(4_c_int32_t, 0_c_int32_t, chaptype == 2)
    desc%sv = c_loc(sv)
    desc%bptnr = c_loc(bptnr)
    pd = trim(param_dir) // c_null_char
    call check(pdmd_ctx_create(pd, ctx), 'pdmd_ctx_create')
    call check(pdmd_set_waves(ctx, int(waves, c_int32_t)), 'pdmd_set_waves')
    call check(pdmd_system_add(ctx, desc, sid), 'pdmd_system_add')
    write(6, '(a,a)') ' engine ', f_string(pdmd_version())
    ! main.F90:345-350 and 425: the starting configuration must be free of
    ! overlaps and bond violations
    call check(pdmd_check(ctx, sid, c_null_ptr, chk), 'pdmd_check')
    if (chk%overlaps + chk%bonds_long + chk%bonds_short > 0) then
      write(6, *) 'found overlap/underlap in initial configuration'
      write(6, *) 'particles ', chk%first_i, ' and ', chk%first_j
      call exit(-1)
    end if
    boundbad = chk%boundbad
    unboundbad = chk%unboundbad

    open(newunit=ue, file='results/run' // digits // '.energy', status='new')
    call check(pdmd_stats(ctx, sid, st), 'pdmd_stats')
    call observables_line(0_c_int64_t)
    coll = 0
    ! main loop: commit entries up to each output pseudo-event, then write
    ! what main.F90:1191-1246 writes there
    do while (coll < ncoll)
      before = st%coll
      left = ncoll - coll
      call check(pdmd_advance_to_output(ctx, 1_c_int32_t, [sid], [left]), 'pdmd_advance_to_output')
      call check(pdmd_stats(ctx, sid, st), 'pdmd_stats')
      if (st%status /= 0) call die('device status error: ' // pdmd_errmsg())
      coll = coll + (st%coll - before)
      if (st%coll - before < left) then
        call observables_line(coll)
        call write_config(coll)
        if (debugging /= 0) call debug_checks(coll)
      end if
    end do
    call observables_line(coll - 1)  ! main.F90:1303
    call write_config(coll)
    close(ue)
    ! main.F90:1320-1322: the end configuration is checked too
    call check(pdmd_check(ctx, sid, c_null_ptr, chk), 'pdmd_check')
    if (chk%overlaps + chk%bonds_long + chk%bonds_short > 0) &
      write(6, *) 'overlap/underlap in the end configuration: particles ', chk%first_i, ' and ', chk%first_j
    call check(pdmd_observables(ctx, sid, ob), 'pdmd_observables')
    write(6, *) ' '
    write(6, *) 'at coll', coll
    write(6, *) 'the final total energy of system (e) is ', ob%ered
    write(6, *) 'the final temperature of system (kt) is ', ob%tred
    write(6, *) 'the final number of alpha-helical hb', ob%hb_alpha
    write(6, *) 'the final number of hydrogen bonds', ob%hb_ii + ob%hb_ij
    write(6, *) 'number of ghost collisions', st%numghosts, ' neighbour-list updates', st%nupdates
    do k = 1, 30
      if (st%nevents(k) > 0) write(6, '(a,i3,a,i12)') ' events of type ', k, ': ', st%nevents(k)
    end do
    call check(pdmd_ctx_destroy(ctx), 'pdmd_ctx_destroy')

  contains

    ! main.F90:1194-1204 (-Ddebugging): overlaps stop the run; changes of
    ! the H-bond consistency counts are reported
    subroutine debug_checks(at)
      integer(c_int64_t), intent(in) :: at
      call check(pdmd_check(ctx, sid, c_null_ptr, chk), 'pdmd_check')
      if (chk%overlaps + chk%bonds_long + chk%bonds_short > 0) then
        write(6, *) 'found overlap/underlap at coll', at
        write(6, *) 'particles ', chk%first_i, ' and ', chk%first_j
        call exit(-1)
      end if
      if (chk%boundbad /= boundbad .or. chk%unboundbad /= unboundbad) then
        write(6, *) 'boundbad and unboundbad', chk%boundbad, chk%unboundbad
        boundbad = chk%boundbad
        unboundbad = chk%unboundbad
      end if
      if (chk%ss_counted /= chk%ss_total) then
        write(6, *) 'wrong ss at coll', at, chk%ss_counted, chk%ss_total
        call exit(-1)
      end if
    end subroutine debug_checks

    subroutine codes_out()
      integer(c_int32_t), allocatable :: a(:), b(:), cd(:)
      integer :: m, q, n, uo
      desc%nchains = int(nch, c_int32_t)
      desc%chain_len = c_loc(clen)
      desc%species = c_loc(spec)
      desc%restype = c_loc(rt)
      desc%hp = c_loc(hp)
      desc%boxl = boxl
      desc%tstar = tstar
      desc%seed = seed
      desc%canon = int(canon, c_int32_t)
      desc%list_capacity = 0
      desc%flags = merge(1_c_int32_t, 0_c_int32_t, no_hbs /= 0) + merge(2_c_int32_t, 0_c_int32_t, no_h /= 0) &
                 + merge(4_c_int32_t, 0_c_int32_t, chaptype == 2)
      desc%sv = c_loc(sv)
      desc%bptnr = c_null_ptr
      ! the first two chains: same-chain and inter-chain pairs
      m = N
      if (nch >= 2) m = min(N, 4 * (clen(1) + clen(2)))
      n = m * (m - 1) / 2
      allocate(a(n), b(n), cd(n))
      n = 0
      do k = 1, m - 1
        do q = k + 1, m
          n = n + 1
          a(n) = k
          b(n) = q
        end do
      end do
      call check(pdmd_pair_codes(desc, int(n, c_int32_t), a, b, cd), 'pdmd_pair_codes')
      write(6, '(a,i8)') ' bytes on the device ', pdmd_system_bytes(desc)
      open(newunit=uo, file='codes.txt', status='replace')
      do k = 1, n
        write(uo, '(3i8)') a(k), b(k), cd(k)
      end do
      close(uo)
    end subroutine codes_out

    subroutine observables_line(c_out)
      integer(c_int64_t), intent(in) :: c_out
      call check(pdmd_observables(ctx, sid, ob), 'pdmd_observables')
      call check(pdmd_stats(ctx, sid, st), 'pdmd_stats')
      tsim = (st%t + st%tfalse) * sqrt(setemp) / sigma1
      write(ue, 22223) c_out, tsim, ob%ered, ob%tred, ob%hb_alpha, ob%hb_ii, ob%hb_ij, ob%ehh_ii, ob%ehh_ij, &
                       ob%rg_avg, ob%e2e_avg
      flush(ue)
22223 format(i15, 3f12.4, 3i8, 4f12.4)
    end subroutine observables_line

    ! config.f:22-39
    subroutine write_config(c_out)
      integer(c_int64_t), intent(in) :: c_out
      integer :: uc
      call check(pdmd_download(ctx, sid, c_loc(sv), c_null_ptr, c_loc(bptnr), c_null_ptr, c_null_ptr, c_null_ptr, &
                               c_null_ptr), 'pdmd_download')
      call check(pdmd_stats(ctx, sid, st), 'pdmd_stats')
      do k = 1, N
        x(k) = sv(6 * (k - 1) + 1) + sv(6 * (k - 1) + 4) * st%tfalse
        y(k) = sv(6 * (k - 1) + 2) + sv(6 * (k - 1) + 5) * st%tfalse
        z(k) = sv(6 * (k - 1) + 3) + sv(6 * (k - 1) + 6) * st%tfalse
      end do
      open(newunit=uc, file='results/run' // digits // '.config', status='unknown', form='unformatted', &
           position='append')
      write(uc) c_out, st%t + st%tfalse, x, y, z
      close(uc)
      open(newunit=uc, file='results/run' // digits // '.bptnr', status='unknown', form='unformatted', &
           position='append')
      write(uc) c_out, bptnr
      close(uc)
      do k = 1, N
        x(k) = sv(6 * (k - 1) + 4)
        y(k) = sv(6 * (k - 1) + 5)
        z(k) = sv(6 * (k - 1) + 6)
      end do
      open(newunit=uc, file='results/run' // digits // '.lastvel', status='replace', form='unformatted')
      write(uc) c_out, x, y, z
      close(uc)
    end subroutine write_config

  end subroutine run

end module pdmd_driver_m

program pdmd_driver
  use pdmd_driver_m
  implicit none
  read(5, *) tstar
  read(5, *) ncoll
  setemp = tstar * 12.d0  ! main.F90:127
  open(newunit=u, file='pdmd.nml', status='old', iostat=ios)
  if (ios == 0) then
    read(u, nml=pdmd_run)
    close(u)
  end if
  if (trim(mode) == 'generate') then
    call generate_start()
  else
    call run()
  end if
end program pdmd_driver
