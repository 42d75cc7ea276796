! ref_harness.F90 -- test infrastructure: drives the reference's own
! subroutines (compiled from their unmodified sources, pulled in below with
! #include through symlinks that oracle/ref/Makefile makes) to produce
! golden vectors for the CPU oracle.  Nothing here is shipped or measured.
!
! stdin: line 1 T* (as temp_XXX line 1, main.F90:124-127), line 2 the number
! of calendar entries of the serial loop below, line 3 (optional, default 1)
! the seed of the project RNG that stands in for IFPORT drandm in the
! terminal-bead capture draw (main.F90:1508-1515).  Reads the reference's
! input files from the working directory (parameters/, parametersep/,
! helix.inp start via -Drunh, inputinfo.f:28-66) and writes raw
! little-endian stream files:
!   setup.bin  scaled tables, bond tables, masses, cell geometry, sv
!   evcode.bin ev_code(N,N) as int32, column-major
!   nb.bin     na_npt(N), nnabdn(N), then each bead's up list and down list
!   cal0.bin   tim(N+3), nptnr(N+3), coltype(N+3) after events()
!   pred.bin   for every pair i<j (row-major in i): tij, type as events.f
!              dispatches them (events.f:26-47)
!   loop.bin   per committed entry (i, j, resolved coltype, tfalse)
!   loopn.bin  entries, pair events, neighbour-list rebuilds, forced
!              rebuilds, capture draws, then the tally of resolved coltypes
!              1..30 (main.F90:926)
!   final.bin  sv, tim, nptnr, coltype, tfalse, identity, bptnr,
!              extra_repuls after the loop
!   obs.bin    energy/radgyr/end_to_end before and after the loop
!   restore.bin identity, bptnr, extra_repuls, ev_code after the -Drunr
!              restart restore (main.F90:240-321) of the final state
!
! The serial loop commits calendar entries in the reference's order: the
! earliest entry of the first non-empty bin, first in list order on ties
! (main.F90:496-545 with one worker), every bead pair event through the
! worker's resolution (main.F90:1432-1960, restated below around calls of
! the reference's repuls_check, repuls_check_3, check_sigma, eventdyn,
! bumpoff, repuls_add, repuls_del_a) and the master's partial_events call
! (main.F90:943), and the sync and output pseudo-events.  It runs without
! the Andersen thermostat (tim(N+1) = 1e9), as the reference does without
! -Dcanon.

#include "def.h"
#include "header.f"

	program ref_harness
	use global
	implicit none
	integer i, j, k, kk, m, evcode, typ, nsteps, npair, nbrsum, nlog, nupdates_h, nforced_h
	logical update
	real*8 tim_sel
	integer, allocatable :: code32(:)
	integer, allocatable :: coltypes(:)
	real*8 tstar_in, tij, tmin, ered, tred, sumvel, ehh_ii, ehh_ij, rg_avg, e2e_avg
	integer hb_alpha, hb_ii, hb_ij
	logical over
	integer boundbad, unboundbad
	real*8, allocatable :: svk(:,:)
	real*8 ev_rec(2)
	integer ev_int(3)
	integer(8) rng_seed, rng_ctr1
	integer ios, tally(30)
	integer, allocatable :: ident0(:), bptnr_fin(:)
	real*8 rxr, ryr, rzr, diffr
	interface
		real(8) function ref_u01(seed, stream, ctr) bind(C, name='ref_u01')
		integer(8), value :: seed, ctr
		integer(4), value :: stream
		end function
	end interface

	noptotal = nop1 + nop2
	coll = 0
	nupdates_h = 0
	nforced_h = 0
	update = .false.
	fname_digits_pre = '0000'
	read(5,*) tstar_in
	read(5,*) nsteps
	rng_seed = 1
	read(5,*,iostat=ios) rng_seed
	if (ios /= 0) rng_seed = 1
	rng_ctr1 = 0
	tally = 0
	setemp = tstar_in*12.d0
	pi = 4.d0*datan(1.d0)
	! main.F90:144-149
	t_fact = 0.00005D0
	n_forced = 150.0d0
	interval = t_fact/dsqrt(setemp)
	interval_max = n_forced*interval
	sortsize = interval_max/dble(numbin)

	call inputinfo()
	call make_code()
	allocate(ident0(noptotal), bptnr_fin(noptotal))
	ident0(1:noptotal) = identity(1:noptotal)

	! main.F90:196-230
	t = 0.d0
	tfalse = 0.d0
	nbin = 1
	tbin_off = 0.d0
	do k = 1, noptotal
		sv(1,k) = sv(1,k) - dnint(sv(1,k))
		sv(2,k) = sv(2,k) - dnint(sv(2,k))
		sv(3,k) = sv(3,k) - dnint(sv(3,k))
		old_rx(k) = sv(1,k)
		old_ry(k) = sv(2,k)
		old_rz(k) = sv(3,k)
		tim(k) = interval_max + ltstep
		coltype(k) = -1
		nptnr(k) = -1
		npt(k) = (k-1)*maxnbs + 1
		npt_dn(k) = (k-1)*maxnbs + 1
		bptnr(k) = 0
		do kk = 1, 4
			extra_repuls(k,kk) = 0
		end do
	end do
	npt(noptotal+1) = noptotal*maxnbs + 1
	npt_dn(noptotal+1) = noptotal*maxnbs + 1
	do k = noptotal+1, noptotal+3
		nptnr(k) = -2
		coltype(k) = -2
	end do
	do k = 0, numbin+1
		bin(k) = 0
	end do
	do k = 1, noptotal+3
		tlinks(k) = 0
		tlinks2(k) = 0
	end do

	! main.F90:389-424 (no Andersen thermostat: tim(N+1) = 1e9)
	call nbor_setup()
	num_cell = int(boxl/(sig_max_all*rl_const)*n_wrap) + 2*n_wrap
	allocate(cell(num_cell**3+1))
	allocate(wrap_map(num_cell**3))
	width = boxl/dble(num_cell-2*n_wrap)
	half = boxl/2.d0
	call cell_link
	call nbor()
	nbrsum = 0
	do k = 1, noptotal-1
		nbrsum = nbrsum + na_npt(k)
	end do
	tim(noptotal+1) = 1000000000.d0
	tim(noptotal+2) = interval
	tim(noptotal+3) = 3.3/(dsqrt(setemp)) + 5
	call events()

	open(unit=21, file='setup.bin', access='stream', form='unformatted', status='replace')
	write(21) sigma_sq, welldia_sq, ep_sqrt, shlddia_sq, ev_param, sqz610
	write(21) bm(1:noptotal)
	write(21) bdln, bl_rn, bl_rc, del_bdln, del_blrn, del_blrc
	write(21) rlsq, sig_max_all, hdelr, width, dble(num_cell), interval, interval_max, sortsize, setemp
	write(21) sv
	write(21) identity, chnnum, hp
	close(21)

	allocate(code32(noptotal*noptotal))
	do j = 1, noptotal
		do i = 1, noptotal
			code32((j-1)*noptotal + i) = ev_code(i,j)
		end do
	end do
	open(unit=21, file='evcode.bin', access='stream', form='unformatted', status='replace')
	write(21) code32
	close(21)

	open(unit=21, file='nb.bin', access='stream', form='unformatted', status='replace')
	write(21) na_npt, nnabdn
	do i = 1, noptotal
		if (na_npt(i) > 0) write(21) nb(npt(i):npt(i)+na_npt(i)-1)
	end do
	do i = 1, noptotal
		if (nnabdn(i) > 0) write(21) dnnab(npt_dn(i):npt_dn(i)+nnabdn(i)-1)
	end do
	close(21)

	open(unit=21, file='cal0.bin', access='stream', form='unformatted', status='replace')
	write(21) tim, nptnr, coltype
	close(21)

	! every pair, dispatched on ev_code as events.f:26-47
	open(unit=21, file='pred.bin', access='stream', form='unformatted', status='replace')
	allocate(coltypes(noptotal+3))
	coltypes = coltype
	do i = 1, noptotal-1
		do j = i+1, noptotal
			tij = 1000000000.d0
			typ = 0
			evcode = ev_code(j,i)
			if (evcode .le. 3) then
				call core(i,j,evcode,tij,typ)
			elseif ((evcode .ge. 4) .and. (evcode .le. 12)) then
				call bond(i,j,evcode,tij,typ)
			elseif (evcode .eq. 15) then
				call nc_sqwel(i,j,evcode,tij,typ)
			elseif (evcode .eq. 16) then
				call sqwel(i,j,evcode,tij,typ)
			elseif ((evcode .ge. 17) .and. (evcode .le. 26)) then
				call core(i,j,evcode,tij,typ)
			elseif (evcode .ge. 40) then
				call sqshlder(i,j,evcode,tij,typ)
			endif
			write(21) tij, typ
		end do
	end do
	close(21)
	coltype = coltypes

	call energy(ered,tred,sumvel,hb_alpha,hb_ii,hb_ij,ehh_ii,ehh_ij)
	call radgyr(rg_avg)
	call end_to_end(e2e_avg)
	open(unit=22, file='obs.bin', access='stream', form='unformatted', status='replace')
	write(22) ered, tred, sumvel, ehh_ii, ehh_ij, rg_avg, e2e_avg, hb_alpha, hb_ii, hb_ij

	! serial loop: selection as main.F90:496-545 with one worker (earliest
	! entry of the first non-empty bin, first in list order on ties), bead
	! pair events through resolve_pair below, the sync pseudo-event as
	! main.F90:1126-1187 and the output pseudo-event's reschedule as
	! main.F90:1238-1242
	open(unit=21, file='loop.bin', access='stream', form='unformatted', status='replace')
	npair = 0
	nlog = 0
	do m = 1, nsteps
		do while (bin(nbin) == 0)
			nbin = nbin + 1
		end do
		i = bin(nbin)
		k = tlinks(i)
		do while (k .ne. 0)
			if (tim(k) < tim(i)) i = k
			k = tlinks(k)
		end do
		j = 0
		if (i == noptotal+1) exit
		tim_sel = tim(i)
		tfalse = tim(i)
		coll = coll + 1
		if (i <= noptotal) then
			j = nptnr(i)
			call resolve_pair(i, j, typ)
			if (typ >= 1 .and. typ <= 30) tally(typ) = tally(typ) + 1
			npair = npair + 1
		else if (i == noptotal+2) then
			typ = 0
			t = t + tfalse
			do k = 1, noptotal+3
				tim(k) = tim(k) - tfalse
			end do
			interval_max = interval_max - tfalse
			tbin_off = tbin_off + tfalse
			do k = 1, noptotal
				sv(1,k) = sv(1,k) + sv(4,k)*tfalse
				sv(2,k) = sv(2,k) + sv(5,k)*tfalse
				sv(3,k) = sv(3,k) + sv(6,k)*tfalse
			end do
			tfalse = 0.d0
			call displ(update)
			if ((update) .or. (interval .gt. interval_max)) then
				if (.not. update) then
					nforced_h = nforced_h + 1
					n_forced = n_forced * 1.01d0
				endif
				interval_max = interval * n_forced
				sortsize = interval_max/dble(numbin)
				tbin_off = 0.d0
				update = .false.
				nupdates_h = nupdates_h + 1
				do k = 1, noptotal
					sv(1,k) = sv(1,k) - anint(sv(1,k))
					sv(2,k) = sv(2,k) - anint(sv(2,k))
					sv(3,k) = sv(3,k) - anint(sv(3,k))
					old_rx(k) = sv(1,k)
					old_ry(k) = sv(2,k)
					old_rz(k) = sv(3,k)
				end do
				call nbor()
				do k = 1, noptotal
					tim(k) = interval_max + ltstep
					coltype(k) = -1
					nptnr(k) = -1
				end do
				call events()
				nbin = 1
			endif
			if (tlinks2(i) .ne. 0) call del_tbin(i)
			tim(i) = interval*0.999d0
			call add_tbin(i)
		else
			typ = 0
			if (tlinks2(i) .ne. 0) call del_tbin(i)
			tim(i) = 3.3d0/(sqrt(setemp)) + 5 + tfalse
			if (tim(i) .lt. interval_max) call add_tbin(i)
		endif
		nlog = nlog + 1
		ev_int(1) = i
		ev_int(2) = j
		ev_int(3) = typ
		ev_rec(1) = tim_sel
		write(21) ev_int, ev_rec(1)
	end do
	close(21)
	open(unit=21, file='loopn.bin', access='stream', form='unformatted', status='replace')
	write(21) nlog, npair, nupdates_h, nforced_h, int(rng_ctr1)
	write(21) tally
	close(21)

	open(unit=21, file='final.bin', access='stream', form='unformatted', status='replace')
	write(21) sv, tim, nptnr, coltype, tfalse
	write(21) identity, bptnr, extra_repuls
	close(21)
	bptnr_fin(1:noptotal) = bptnr(1:noptotal)

	call energy(ered,tred,sumvel,hb_alpha,hb_ii,hb_ij,ehh_ii,ehh_ij)
	call radgyr(rg_avg)
	call end_to_end(e2e_avg)
	write(22) ered, tred, sumvel, ehh_ii, ehh_ij, rg_avg, e2e_avg, hb_alpha, hb_ii, hb_ij
	close(22)

	! checkover.f and check_nc_int.f on the final state, then checkover.f on
	! two planted states (bead 5 on top of bead noptotal-3; bead 10 moved by
	! 0.05 box lengths along x); their messages go to stdout between the
	! CHECK and CHECKEND lines
	write(6,'(a)') 'CHECK final'
	call checkover(over)
	boundbad = 0
	unboundbad = 0
	call check_nc_int(boundbad, unboundbad)
	write(6,'(a,l2,2i8)') 'CHECKEND ', over, boundbad, unboundbad
	allocate(svk(6, noptotal))
	svk(:, 1:noptotal) = sv(:, 1:noptotal)
	sv(1,5) = sv(1,noptotal-3) + 1.d-3
	sv(2,5) = sv(2,noptotal-3) + 1.d-3
	sv(3,5) = sv(3,noptotal-3) + 1.d-3
	write(6,'(a)') 'CHECK planted1'
	call checkover(over)
	write(6,'(a,l2)') 'CHECKEND ', over
	sv(:, 1:noptotal) = svk(:, 1:noptotal)
	sv(1,10) = sv(1,10) + 5.d-2
	write(6,'(a)') 'CHECK planted2'
	call checkover(over)
	write(6,'(a,l2)') 'CHECKEND ', over
	sv(:, 1:noptotal) = svk(:, 1:noptotal)

	! -Drunr restart from the final state: the positions as a restart reads
	! them back (wrapped by inputinfo.f:90-95 and again by main.F90:210-216),
	! identities and codes from the input files (inputinfo.f, make_code.f),
	! bptnr from the .bptnr record, then the restore of main.F90:240-321:
	! every pair k < kj in order, an N-C pair inside its well gets its
	! shoulders (repuls_add) unless an end bead takes part, and the pair
	! (k, bptnr(k)) takes the bound identities 5/8.  check_nc_int.f then
	! checks the restored state on fresh lists.
	do k = 1, noptotal
		do kk = 1, 3
			sv(kk,k) = sv(kk,k) - dnint(sv(kk,k))
			sv(kk,k) = sv(kk,k) - dnint(sv(kk,k))
		end do
		identity(k) = ident0(k)
		bptnr(k) = bptnr_fin(k)
		do kk = 1, 4
			extra_repuls(k,kk) = 0
		end do
	end do
	call make_code()
	do k = 1, noptotal-1
		do kk = k+1, noptotal
			if (identity(k)+identity(kk) .eq. 5 .and. ev_code(k,kk) .eq. 15) then
				rxr = sv(1,k) - sv(1,kk)
				ryr = sv(2,k) - sv(2,kk)
				rzr = sv(3,k) - sv(3,kk)
				rxr = rxr - dnint(rxr)
				ryr = ryr - dnint(ryr)
				rzr = rzr - dnint(rzr)
				diffr = rxr*rxr + ryr*ryr + rzr*rzr - welldia_sq(identity(k),identity(kk))
				if (diffr .lt. 0.d0 .and. .not. is_end(k) .and. .not. is_end(kk)) then
					if (identity(k) .eq. 1) then
						call repuls_add(k,kk)
					else
						call repuls_add(kk,k)
					endif
				endif
			endif
			if (kk .eq. bptnr(k)) then
				if (identity(k) .eq. 1) then
					identity(k) = 5
					identity(kk) = 8
				else
					identity(k) = 8
					identity(kk) = 5
				endif
			endif
		end do
	end do
	tfalse = 0.d0
	do k = 1, noptotal
		old_rx(k) = sv(1,k)
		old_ry(k) = sv(2,k)
		old_rz(k) = sv(3,k)
	end do
	call nbor()
	write(6,'(a)') 'CHECK restored'
	boundbad = 0
	unboundbad = 0
	call check_nc_int(boundbad, unboundbad)
	write(6,'(a,l2,2i8)') 'CHECKEND ', .false., boundbad, unboundbad
	do j = 1, noptotal
		do i = 1, noptotal
			code32((j-1)*noptotal + i) = ev_code(i,j)
		end do
	end do
	open(unit=21, file='restore.bin', access='stream', form='unformatted', status='replace')
	write(21) identity, bptnr, extra_repuls, code32
	close(21)

	write(6,'(a,i8,a,f10.4)') 'harness: pair events committed ', npair, ' avg nbors ', dble(nbrsum)/dble(noptotal)

	contains

	! end bead of its chain as main.F90:1483-1494 and 1493-1494 test it:
	! the N bead of the first residue or the C bead of the last
	logical function is_end(k)
	integer, intent(in) :: k
	integer kl
	if (k .le. nop1) then
		kl = k - (chnnum(k)-1)*numbeads1
		is_end = (kl .eq. chnln1+1) .or. (kl .eq. 3*chnln1)
	else
		kl = k - nop1 - (chnnum(k)-(nop1/numbeads1)-1)*numbeads2 + numbeads1
		is_end = (kl .eq. numbeads1+chnln2+1) .or. (kl .eq. numbeads1+3*chnln2)
	endif
	end function is_end

	! One bead pair event as a worker resolves it (main.F90:1432-1960) on
	! the current state, then the master's commit (main.F90:641-943, whose
	! table writes repeat the worker's in one address space) and
	! partial_events.  The capture draw of an end-bead N-C pair uses the
	! project RNG (stream 1) in place of drandm.
	subroutine resolve_pair(i, j, typ)
	integer, intent(in) :: i, j
	integer, intent(out) :: typ
	integer evc, hbp, x
	logical ends, xdel
	real*8 rating, u
	evc = ev_code(i,j)
	ends = is_end(i) .or. is_end(j)
	xdel = .false.
	if (coltype(i) .eq. 7) then
		if ((extra_repuls(i,4) .eq. 0) .and. (extra_repuls(j,4) .eq. 0)) then
			if (.not. ends) then
				if (identity(i) < identity(j)) then
					call repuls_check(i,j,rating)
				else
					call repuls_check(j,i,rating)
				endif
				if (rating .le. 10.d0) then
					coltype(i) = 4
				else
					coltype(i) = 14
				endif
			else
				u = ref_u01(rng_seed, 1, rng_ctr1)
				rng_ctr1 = rng_ctr1 + 1
				if (u .le. 0.2d0) then
					coltype(i) = 4
				else
					coltype(i) = 9
				endif
			endif
		else
			coltype(i) = 9
		endif
	elseif (coltype(i) .eq. 10) then
		if (evc < 45) then
			hbp = extra_repuls(i,4)
			if (identity(i) < identity(hbp)) then
				call repuls_check_3(i,hbp,j,rating)
			else
				call repuls_check_3(hbp,i,j,rating)
			endif
		else
			hbp = extra_repuls(j,4)
			if (identity(j) < identity(hbp)) then
				call repuls_check_3(j,hbp,i,rating)
			else
				call repuls_check_3(hbp,j,i,rating)
			endif
		endif
		if (rating .le. 10.d0) then
			coltype(i) = 5
		else
			coltype(i) = 15
		endif
	elseif (coltype(i) .eq. 12) then
		if (evc < 45) then
			x = i
		else
			x = j
		endif
		hbp = extra_repuls(x,4)
		if (bptnr(x) == hbp) then
			call check_sigma(x,hbp,rating)
			if (rating .gt. 10.d0) then
				coltype(i) = 6
			else
				coltype(i) = 13
			endif
		else
			coltype(i) = 15
		endif
	endif
	if (coltype(i) .lt. 14) call eventdyn(i,j,evc)
	if (coltype(i) .eq. 20) then
		if (identity(i)+identity(j) .eq. 5) then
			bptnr(i) = j
			bptnr(j) = i
			identity(i) = identity(i)+4
			identity(j) = identity(j)+4
			if (.not. ends) then
				if (identity(i) .lt. identity(j)) then
					call repuls_add(i,j)
				else
					call repuls_add(j,i)
				endif
			endif
		endif
	elseif (coltype(i) .eq. 21) then
		if (identity(i) .le. 8) then
			if (.not. ends) then
				if (identity(i) .lt. identity(j)) then
					call repuls_del_a(i,j)
				else
					call repuls_del_a(j,i)
				endif
				xdel = .true.
			endif
			if (bptnr(i) == j) then
				bptnr(i) = 0
				bptnr(j) = 0
				identity(i) = identity(i)-4
				identity(j) = identity(j)-4
			endif
		endif
	elseif (coltype(i) .eq. 24) then
		if (evc < 45) then
			x = i
		else
			x = j
		endif
		hbp = extra_repuls(x,4)
		if (identity(x) + identity(hbp) .eq. 5) then
			bptnr(x) = hbp
			bptnr(hbp) = x
			identity(x) = identity(x)+4
			identity(hbp) = identity(hbp)+4
		endif
	elseif (coltype(i) .eq. 25) then
		if (evc < 45) then
			x = i
		else
			x = j
		endif
		hbp = extra_repuls(x,4)
		if (identity(x) .ge. 5) then
			bptnr(x) = 0
			bptnr(hbp) = 0
			identity(x) = identity(x)-4
			identity(hbp) = identity(hbp)-4
		endif
	elseif (coltype(i) .eq. 14) then
		call bumpoff(i,j,evc)
		if (.not. ends) then
			if (identity(i) .lt. identity(j)) then
				call repuls_add(i,j)
			else
				call repuls_add(j,i)
			endif
		endif
	elseif (coltype(i) .eq. 15) then
		call bumpoff(i,j,evc)
	elseif (coltype(i) .eq. 16) then
		call bumpoff(i,j,evc)
		if (.not. ends) then
			if (identity(i) .lt. identity(j)) then
				call repuls_del_a(i,j)
			else
				call repuls_del_a(j,i)
			endif
			xdel = .true.
		endif
	endif
	typ = coltype(i)
	call partial_events(i,j,xdel)
	end subroutine resolve_pair

	end program ref_harness

#include "add_tbin.f"
#include "del_tbin.f"
#include "bond.f"
#include "bumped.f"
#include "cell_add.f"
#include "cell_link.f"
#include "core.f"
#include "displ.f"
#include "energy.f"
#include "eventdyn.f"
#include "eventredo_up.f"
#include "eventredo_down.f"
#include "events.f"
#include "inputinfo.f"
#include "make_code.f"
#include "nbor.f"
#include "nbor_setup.f"
#include "nc_sqwel.f"
#include "partial_events.f"
#include "repuls_add.f"
#include "repuls_del_a.f"
#include "repuls_del_b.f"
#include "scale_down.f"
#include "sqshlder.f"
#include "sqwel.f"
#include "radgyr.f"
#include "end2end.f"
#include "repuls_check.f"
#include "repuls_check_3.f"
#include "check_sigma.f"
#include "checkover.f"
#include "check_nc_int.f"
