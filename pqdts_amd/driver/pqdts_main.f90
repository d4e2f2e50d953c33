! pqdts_omp.x drop-in: host driver of the MI355X solver.
!
! Same command line (14 positional arguments, pqdts.f90:1243-1270; README.md
! "Command Line Arguments"), same inputs (data/F_*.bin, data/P_*.bin, optional
! data/initial_     0.dat) and same outputs (rank_     0_oiter     k.dat,
! rank_     0_final.dat, timing_           0 .out, status_           0 .out)
! as the reference program, so pqdts.py drives it unchanged. The numerical
! work runs on one GPU behind the C ABI of include/pqdts_amd.h, bound here
! through iso_c_binding. stdout uses explicit formats that reproduce the
! fields pqdts.py parses (a line starting with " iter   " whose fifth
! whitespace token is the objective, pqdts.py:77-78).
!
! PQDTS_POLISH=k (environment; no counterpart in the reference) continues the
! solve after the reference's stopping test with at most k polish iterations
! to the problem's minimiser (pqdts_set_polish). The per-stage files keep the
! reference's stopping point; rank_*_final.dat holds the polished result. Its
! progress lines start with " polish", which pqdts.py's parse ignores.
! PQDTS_POLISH_METHOD=1 with it selects the direct face solves of the polish
! (pqdts_set_polish_method); 0, the default, its Newton-CG steps.
module pqdts_capi
  use iso_c_binding
  implicit none
  integer(c_int), parameter :: SLOT_X = 0, SLOT_G = 1, SLOT_HP = 2, SLOT_XTMP = 3, SLOT_SN = 4, SLOT_Z = 5
  integer(c_int), parameter :: COL_MAJOR = 1
  integer(c_int), parameter :: EV_ITER = 0, EV_PCG = 1, EV_LS = 2, EV_ACTIVE = 3, EV_STOP1 = 4, &
                               EV_ITER0 = 5, EV_LAMBDA = 6, EV_STAGE = 7, EV_POLISH = 8, &
                               EV_POLISHED = 9
  interface
    function pqdts_last_error() bind(C, name="pqdts_last_error")
      import :: c_ptr
      type(c_ptr) :: pqdts_last_error
    end function
    integer(c_int) function pqdts_device_count() bind(C, name="pqdts_device_count")
      import :: c_int
    end function
    integer(c_int) function pqdts_create(M, N, D, nF, Fr, Fc, Fv, nP, Pr, Pc, Pv, gamma, ctx) &
      bind(C, name="pqdts_create")
      import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
      integer(c_int64_t), value :: M, N, D, nF, nP
      integer(c_int32_t) :: Fr(*), Fc(*), Pr(*), Pc(*)
      real(c_double) :: Fv(*), Pv(*)
      real(c_double), value :: gamma
      type(c_ptr) :: ctx
    end function
    integer(c_int) function pqdts_destroy(ctx) bind(C, name="pqdts_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function pqdts_put(ctx, slot, host, layout) bind(C, name="pqdts_put")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: slot, layout
      real(c_double) :: host(*)
    end function
    integer(c_int) function pqdts_get(ctx, slot, host, layout) bind(C, name="pqdts_get")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: slot, layout
      real(c_double) :: host(*)
    end function
    integer(c_int) function pqdts_set_maxr(ctx, maxr) bind(C, name="pqdts_set_maxr")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t) :: maxr(*)
    end function
    integer(c_int) function pqdts_L(ctx, x, val) bind(C, name="pqdts_L")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: x
      real(c_double) :: val
    end function
    integer(c_int) function pqdts_dL(ctx, x, g) bind(C, name="pqdts_dL")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: x, g
    end function
    integer(c_int) function pqdts_BF(ctx, dd, bx) bind(C, name="pqdts_BF")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: dd, bx
    end function
    integer(c_int) function pqdts_Bdiag(ctx, diag) bind(C, name="pqdts_Bdiag")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: diag(*)
    end function
    integer(c_int) function pqdts_doproj(ctx, x) bind(C, name="pqdts_doproj")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: x
    end function
    integer(c_int) function pqdts_pdot(ctx, x, y, val) bind(C, name="pqdts_pdot")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: x, y
      real(c_double) :: val
    end function
    integer(c_int) function pqdts_paxpy(ctx, a, x, y) bind(C, name="pqdts_paxpy")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: a
      integer(c_int), value :: x, y
    end function
    integer(c_int) function pqdts_pgaxpy(ctx, a, x, b, y) bind(C, name="pqdts_pgaxpy")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: a, b
      integer(c_int), value :: x, y
    end function
    integer(c_int) function pqdts_pgaxpy2(ctx, a, x, b, y, z) bind(C, name="pqdts_pgaxpy2")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: a, b
      integer(c_int), value :: x, y, z
    end function
    integer(c_int) function pqdts_pset(ctx, a, x) bind(C, name="pqdts_pset")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: a
      integer(c_int), value :: x
    end function
    integer(c_int) function pqdts_pcopy(ctx, x, y) bind(C, name="pqdts_pcopy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: x, y
    end function
    integer(c_int) function pqdts_pscale(ctx, a, x) bind(C, name="pqdts_pscale")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: a
      integer(c_int), value :: x
    end function
    integer(c_int) function pqdts_pcopy_scale(ctx, a, x, y) bind(C, name="pqdts_pcopy_scale")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: a
      integer(c_int), value :: x, y
    end function
    integer(c_int) function pqdts_solve(ctx, mode, maxiter, eps, start_stage, output, cb, user, conv) &
      bind(C, name="pqdts_solve")
      import :: c_int, c_ptr, c_double, c_funptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: mode, maxiter, start_stage, output
      real(c_double), value :: eps
      type(c_funptr), value :: cb
      type(c_ptr), value :: user
      real(c_double) :: conv
    end function
    integer(c_int) function pqdts_set_polish(ctx, max_outer, cg_rtol, tol, budget, fp32) &
      bind(C, name="pqdts_set_polish")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: max_outer, fp32
      real(c_double), value :: cg_rtol, tol, budget
    end function
    integer(c_int) function pqdts_set_polish_method(ctx, method) bind(C, name="pqdts_set_polish_method")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: method
    end function
    integer(c_int) function pqdts_device_timers(ctx, on) bind(C, name="pqdts_device_timers")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: on
    end function
    integer(c_int) function pqdts_timers(ctx, cap, names, counts, seconds) bind(C, name="pqdts_timers")
      import :: c_int, c_ptr, c_double, c_char
      type(c_ptr), value :: ctx
      integer(c_int), value :: cap
      character(kind=c_char) :: names(24, *)
      integer(c_int) :: counts(*)
      real(c_double) :: seconds(*)
    end function
  end interface

contains

  function last_error() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: p(:)
    type(c_ptr) :: cp
    integer :: n
    cp = pqdts_last_error()
    call c_f_pointer(cp, p, [4096])
    n = 0
    do while (n < 4096)
      if (p(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate (character(len=n) :: msg)
    msg = transfer(p(1:n), msg)
  end function

  ! Stops like the reference does on a fatal condition (message, exit status 0).
  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    if (rc /= 0) then
      print '(a)', " pqdts_amd error in "//what//": "//last_error()
      stop
    end if
  end subroutine
end module

module driver_state
  use iso_c_binding
  use iso_fortran_env, only: int32, int64, real64
  implicit none
  type(c_ptr) :: ctx = c_null_ptr
  integer(int64) :: M, N, D
  integer(int32) :: output = 0, rank = 0, mode = 2
  real(real64), allocatable :: X(:, :)
  real(real64) :: t0 = 0, tlast = 0
  ! driver-side timers (prep, input, output); the solver's own come from pqdts_timers
  character(len=20) :: tname(8)
  integer(int32) :: tcount(8) = 0
  real(real64) :: tsum(8) = 0, tbeg(8) = 0
  integer :: ntimers = 0
  logical :: timing_open = .false.
contains

  real(real64) function wtime()
    integer(int64) :: cnt, rate
    call system_clock(cnt, rate)
    wtime = real(cnt, real64)/real(rate, real64)
  end function

  integer function tid(name)
    character(*), intent(in) :: name
    integer :: i
    do i = 1, ntimers
      if (trim(tname(i)) == name) then
        tid = i
        return
      end if
    end do
    ntimers = ntimers + 1
    tname(ntimers) = name
    tid = ntimers
  end function

  subroutine tstart(name)
    character(*), intent(in) :: name
    integer :: i
    i = tid(name)
    tcount(i) = tcount(i) + 1
    tbeg(i) = wtime()
  end subroutine

  subroutine tstop(name)
    character(*), intent(in) :: name
    integer :: i
    i = tid(name)
    tsum(i) = tsum(i) + (wtime() - tbeg(i))
  end subroutine

  ! "timing_<I12 rank> .out": replaced at iteration 0, appended afterwards
  ! (pqdts.f90:159-186); rows "iter name count seconds".
  subroutine tprint(iter)
    use pqdts_capi
    integer(int32), intent(in) :: iter
    character(len=64) :: fname
    character(kind=c_char) :: names(24, 64)
    integer(c_int) :: counts(64), n, i, k
    real(c_double) :: secs(64)
    character(len=24) :: nm
    write (fname, '(a,i12,a)') "timing_", rank, " .out"
    if (.not. timing_open) then
      if (iter == 0) then
        open (43, file=trim(fname), status="replace", action="write")
      else
        open (43, file=trim(fname), position="append", action="write")
      end if
      timing_open = .true.
    end if
    do i = 1, ntimers
      write (43, '(i12,1x,a20,i12,es26.16e3)') iter, tname(i), tcount(i), tsum(i)
    end do
    if (c_associated(ctx)) then
      n = pqdts_timers(ctx, 64, names, counts, secs)
      do i = 1, min(n, 64)
        nm = ""
        do k = 1, 24
          if (names(k, i) == c_null_char) exit
          nm(k:k) = names(k, i)
        end do
        write (43, '(i12,1x,a20,i12,es26.16e3)') iter, nm(1:20), counts(i), secs(i)
      end do
    end if
    flush (43)
  end subroutine

  subroutine write_state(fname)
    use pqdts_capi
    character(*), intent(in) :: fname
    integer :: u
    call tstart("output")
    call check(pqdts_get(ctx, SLOT_X, X, COL_MAJOR), "pqdts_get")
    print '(1x,a)', fname
    open (newunit=u, file=fname, status="replace", form="unformatted", access="sequential")
    write (u) M
    write (u) N
    write (u) D
    write (u) M
    write (u) rank
    write (u) X
    close (u)
    call tstop("output")
  end subroutine

  ! Progress events of pqdts_solve, printed in the reference's order and words.
  subroutine on_event(user, kind, values, nvalues) bind(C)
    use pqdts_capi
    type(c_ptr), value :: user
    integer(c_int), value :: kind, nvalues
    type(c_ptr), value :: values
    real(c_double), pointer :: v(:)
    character(len=64) :: fname
    real(real64) :: now
    call c_f_pointer(values, v, [max(nvalues, 1)])
    select case (kind)
    case (EV_ITER0)
      now = wtime()
      print '(a,2i12,a,es25.16e3,a,es25.16e3,a,es25.16e3,a,es25.16e3)', " iter", 0, 0, " L= ", v(1), &
        " ||der||= ", v(2), " dt= ", 0.0_real64, " ttot= ", now - t0
      tlast = now
      call tprint(0_int32)
    case (EV_ACTIVE)
      print '(a,i12,a,i12,es25.16e3)', " positivity constraints", nint(v(1), int64), " active of", &
        nint(v(2), int64), v(1)/v(2)
    case (EV_PCG)
      print '(a,i12,a,es25.16e3,a,es25.16e3,a,es25.16e3,a,es25.16e3)', " newton-pcg", nint(v(1)), &
        " sqrt(b2)= ", v(2), " eps= ", v(3), " b1= ", v(4), " v2= ", v(5)
    case (EV_LS)
      if (nvalues >= 6 .and. v(6) /= 0) then
        print '(a,i12,a,es25.16e3,a,es25.16e3,a,es25.16e3,a,es25.16e3)', " line search", nint(v(1)), &
          " L= ", v(2), " alpha= ", v(3), " df(0)= ", v(4), " min(X)= ", v(5)
      else
        print '(a,i12,a,es25.16e3,a,es25.16e3,a,es25.16e3)', " line search", nint(v(1)), &
          " L= ", v(2), " alpha= ", v(3), " df(0)= ", v(4)
      end if
    case (EV_STOP1)
      print '(a,es25.16e3)', " stopping inner iteration because dldlpha is small", v(1)
    case (EV_ITER)
      now = wtime()
      print '(a,2i12,a,es25.16e3,a,es25.16e3,a,es25.16e3,a,es25.16e3,a,i12,a,es25.16e3,a,es25.16e3,a,'// &
        'es25.16e3,a,es25.16e3)', " iter", nint(v(1)), nint(v(2)), " L= ", v(3), " max(abs(c)) ", v(4), &
        " conv ", v(5), " ||der||= ", v(6), " mode= ", nint(v(7)), " alpha= ", v(8), " dlalpha= ", v(9), &
        " dt= ", now - tlast, " ttot= ", now - t0
      tlast = now
    case (EV_LAMBDA)
      print '(a,i12,2es25.16e3)', " lambda", nint(v(1)), v(2), v(3)
    case (EV_POLISH)
      print '(a,i12,a,i12,a,i12,a,i12,a,es25.16e3)', " polish", nint(v(1)), " active ", nint(v(2), int64), &
        " changed ", nint(v(3), int64), " cg ", nint(v(4)), " step ", v(5)
    case (EV_POLISHED)
      print '(a,i3,a,i12,a,i12,a,es25.16e3,a,es25.16e3)', " polished status ", nint(v(1)), " iterations ", &
        nint(v(2)), " cg ", nint(v(3)), " conv ", v(4), " L= ", v(5)
    case (EV_STAGE)
      if (output >= 2) then
        write (fname, '(a,i6,a,i6,a)') "rank_", rank, "_oiter", nint(v(1)), ".dat"
        call write_state(trim(fname))
      end if
      call tprint(int(nint(v(1)), int32))
    end select
  end subroutine
end module

program pqdts_amd_driver
  use iso_c_binding
  use iso_fortran_env, only: int8, int32, int64, real64
  use pqdts_capi
  use driver_state
  implicit none
  integer(int64) :: NF, NP, i, j, k, nnzF, nnzP
  integer(int32) :: maxiter, start_stage, read_input, bench, smo_dist, smo_c, rank2
  real(real64) :: gam, conv_thres, smo_fact, val, ts, s
  integer(int64) :: M2, N2, D2, lr2
  integer(int32), allocatable :: Fr(:), Fc(:), Pr(:), Pc(:), maxr0(:)
  real(real64), allocatable :: Fv(:), Pv(:), Y(:, :), diag(:), B(:, :)
  character(len=32) :: arg
  character(len=64) :: fname
  logical :: exists
  integer :: u, it, pm
  real(c_double) :: conv

  t0 = wtime()
  call tstart("prep")
  call get_command_argument(1, arg); read (arg, *) M
  call get_command_argument(2, arg); read (arg, *) N
  call get_command_argument(3, arg); read (arg, *) D
  call get_command_argument(4, arg); read (arg, *) NF
  call get_command_argument(5, arg); read (arg, *) NP
  call get_command_argument(6, arg); read (arg, *) mode
  call get_command_argument(7, arg); read (arg, *) maxiter
  call get_command_argument(8, arg); read (arg, *) output
  call get_command_argument(9, arg); read (arg, *) gam
  call get_command_argument(10, arg); read (arg, *) conv_thres
  call get_command_argument(11, arg); read (arg, *) start_stage
  call get_command_argument(12, arg); read (arg, *) read_input
  call get_command_argument(13, arg); read (arg, *) smo_fact
  call get_command_argument(14, arg); read (arg, *) bench

  print '(a,i12)', " num_proc", 1
  print '(a,i12,a,i12,a,i12)', " M=", M, " N=", N, " D=", D
  print '(a,i12,a,i12,a,i12)', " mode=", mode, " maxiiter=", maxiter, " output=", output
  print '(a,es25.16e3,a,es25.16e3)', " gam=", gam, " smo_fact=", smo_fact
  print '(a,es25.16e3)', " conv_thres=", conv_thres
  if (pqdts_device_count() < 1) then
    print '(a)', " pqdts_amd: no GPU visible"
    stop
  end if

  ! the generator is seeded once, at program start (pqdts.f90:1241); P (when random)
  ! and the rows of the maxr X then draw from it in the reference's order
  call random_seed_fixed()

  ! ---- F: from data/F_*.bin (pqdts.f90:1318-1365) or generated (1366-1466)
  if (NF >= 0) then
    call read_coo("data/F", NF, Fr, Fc, Fv)
    nnzF = NF
    print '(i12,a,es25.16e3)', rank, " sum(F)=", sum(Fv)
  else
    call internal_F(M, D, Fr, Fc, Fv, nnzF)
    if (nnzF < 0) stop
  end if

  ! ---- P: data/P_*.bin (pqdts.f90:1527-1535) or random (1537)
  if (NP >= 0) then
    call read_coo("data/P", NP, Pr, Pc, Pv)
    nnzP = NP
    print '(i12,a,es25.16e3)', rank, " sum(P)=", sum(Pv)
  else
    nnzP = N*D
    allocate (Pr(nnzP), Pc(nnzP), Pv(nnzP))
    ! P(N, D) of the reference in memory order: n fastest, then d
    call random_number(Pv)
    k = 0
    do j = 0, D - 1
      do i = 0, N - 1
        k = k + 1
        Pr(k) = int(j, int32)
        Pc(k) = int(i, int32)
      end do
    end do
  end if

  call check(pqdts_create(M, N, D, nnzF, Fr, Fc, Fv, nnzP, Pr, Pc, Pv, gam, ctx), "pqdts_create")
  ! the solver's timers measure device time, as the reference's measure its own work
  call check(pqdts_device_timers(ctx, 1_c_int), "pqdts_device_timers")
  call get_environment_variable("PQDTS_POLISH", arg, status=it)
  if (it == 0) then
    read (arg, *) it
    if (it > 0) call check(pqdts_set_polish(ctx, int(it, c_int), 1e-6_c_double, 1e-10_c_double, &
                                             0.0_c_double, 0_c_int), "pqdts_set_polish")
    call get_environment_variable("PQDTS_POLISH_METHOD", arg, status=pm)
    if (pm == 0) then
      read (arg, *) pm
      call check(pqdts_set_polish_method(ctx, int(pm, c_int)), "pqdts_set_polish_method")
    end if
  end if
  deallocate (Fr, Fc, Fv, Pr, Pc, Pv)
  allocate (X(M, N))

  ! maxr from a random row-stochastic X, one row at a time as the reference draws it
  ! (pqdts.f90:1565-1575); it only matters when the solve starts in stage 2
  allocate (maxr0(M))
  do i = 1, M
    call random_number(X(i, :))
    X(i, :) = X(i, :)/sum(X(i, :))
    maxr0(i) = int(maxloc(X(i, :), 1) - 1, int32)
  end do
  call check(pqdts_set_maxr(ctx, maxr0), "pqdts_set_maxr")

  if (bench == 1) call quick_benchmark()

  ! ---- initial state (pqdts.f90:1833-1903)
  X = 1.0_real64/real(N, real64)
  write (fname, '(a,i6,a)') "data/initial_", rank, ".dat"
  inquire (file=trim(fname), exist=exists)
  if (exists .and. read_input == 1) then
    print '(a,1x,a)', " reading starting state from", trim(fname)
    call tstart("input")
    open (newunit=u, file=trim(fname), status="old", form="unformatted", access="sequential")
    read (u) M2
    read (u) N2
    read (u) D2
    read (u) lr2
    read (u) rank2
    if (M /= M2) then
      print '(a,i12,a,i12)', " input file in this directory is for a different M", M, " !=", M2
      stop
    end if
    if (N /= N2) then
      print '(a,i12,a,i12)', " input file in this directory is for a different N", N, " !=", N2
      stop
    end if
    if (D /= D2) then
      print '(a,i12,a,i12)', " input file in this directory is for a different D", D, " !=", D2
      stop
    end if
    if (M /= lr2) then
      print '(a,i12,a,i12)', " input file in this directory is for a different localrows", M, " !=", lr2
      stop
    end if
    if (rank /= rank2) then
      print '(a,i12,a,i12)', " input file in this directory is for a different rank", rank, " !=", rank2
      stop
    end if
    read (u) X
    close (u)
    call tstop("input")
    if (smo_fact > 0) then
      ! moving-average smoothing for rows > 100, then row renormalisation (pqdts.f90:1882-1902)
      allocate (Y(M, N))
      Y = X
      do j = 101, M
        smo_dist = int(anint(real(j, real64)/smo_fact), int32)
        do i = 1, N
          smo_c = 0
          s = 0
          do k = max(1_int64, j - smo_dist), min(j + smo_dist, M)
            s = s + X(k, i)
            smo_c = smo_c + 1
          end do
          Y(j, i) = s/real(smo_c, real64)
        end do
        Y(j, :) = Y(j, :)/sum(Y(j, :))
      end do
      X = Y
      deallocate (Y)
    end if
  end if
  call check(pqdts_put(ctx, SLOT_X, X, COL_MAJOR), "pqdts_put")
  call tstop("prep")

  ! ---- minimisation (pqdts.f90:1908-2168); prints come back as events
  tlast = wtime()
  call check(pqdts_solve(ctx, mode, maxiter, conv_thres, start_stage, output, c_funloc(on_event), c_null_ptr, &
                         conv), "pqdts_solve")
  ! the reference prints the iteration-0 line before timing row 0 (pqdts.f90:1914-1918)

  if (output >= 1) then
    write (fname, '(a,i6,a)') "rank_", rank, "_final.dat"
    call write_state(trim(fname))
  end if
  call tprint(-1_int32)
  call write_status()
  it = pqdts_destroy(ctx)

contains

  subroutine random_seed_fixed()
    integer :: nseed
    integer, allocatable :: seed(:)
    call random_seed(size=nseed)
    allocate (seed(nseed))
    seed(:) = 23764
    call random_seed(put=seed)
  end subroutine

  ! data/<name>_row.bin, _col.bin (int32, 0-based) and _data.bin (float64),
  ! raw stream files (pqdts.f90:1080-1107)
  subroutine read_coo(base, nnz, r, c, v)
    character(*), intent(in) :: base
    integer(int64), intent(in) :: nnz
    integer(int32), allocatable, intent(out) :: r(:), c(:)
    real(real64), allocatable, intent(out) :: v(:)
    integer :: uu
    allocate (r(nnz), c(nnz), v(nnz))
    open (newunit=uu, file=base//"_row.bin", status="old", access="stream", form="unformatted")
    read (uu) r
    close (uu)
    open (newunit=uu, file=base//"_col.bin", status="old", access="stream", form="unformatted")
    read (uu) c
    close (uu)
    open (newunit=uu, file=base//"_data.bin", status="old", access="stream", form="unformatted")
    read (uu) v
    close (uu)
  end subroutine

  ! photon-number distribution of a coherent probe with mean `mean`: Poisson
  ! below 50, its normal approximation above (pqdts.f90:188-206)
  real(real64) function fprob(x, mean)
    real(real64), intent(in) :: x, mean
    real(real64), parameter :: pi = 4.0_real64*atan(1.0_real64)
    real(real64) :: y
    if (mean < 50) then
      if (mean == 0 .and. x == 0) then
        fprob = 1.0_real64
      else
        fprob = exp(log(mean)*x - log_gamma(x + 1.0_real64) - mean)
      end if
    else
      y = (x - mean)/sqrt(mean)
      fprob = exp(-0.5_real64*y**2)/sqrt(mean)/sqrt(2*pi)
    end if
  end function

  ! Internal F (argument 4 = -1, pqdts.f90:1366-1466): row d (1-based) has mean
  ! photon number (D-d)^2, entries kept while the pmf is >= 1e-10, M = (D-1)^2.
  subroutine internal_F(Mi, Di, r, c, v, nnz)
    integer(int64), intent(in) :: Mi, Di
    integer(int32), allocatable, intent(out) :: r(:), c(:)
    real(real64), allocatable, intent(out) :: v(:)
    integer(int64), intent(out) :: nnz
    integer(int64), allocatable :: lo(:), hi(:)
    integer(int64) :: dd, jj, mean, pos
    real(real64) :: p
    if (Mi /= (Di - 1)**2) then
      print '(a)', " M is not (D-1)**2"
      nnz = -1
      return
    end if
    allocate (lo(Di), hi(Di))
    do dd = 1, Di
      mean = (Di - dd)**2
      lo(dd) = huge(lo(dd))
      hi(dd) = 0
      do jj = max(1_int64, mean), Mi
        p = fprob(real(jj - 1, real64), real(mean, real64))
        if (p < 1d-10) exit
        lo(dd) = min(jj, lo(dd))
        hi(dd) = max(jj, hi(dd))
      end do
      do jj = max(1_int64, mean), 1, -1
        p = fprob(real(jj - 1, real64), real(mean, real64))
        if (p < 1d-10) exit
        lo(dd) = min(jj, lo(dd))
        hi(dd) = max(jj, hi(dd))
      end do
    end do
    nnz = 0
    do dd = 1, Di
      if (hi(dd) >= lo(dd)) nnz = nnz + hi(dd) - lo(dd) + 1
    end do
    allocate (r(nnz), c(nnz), v(nnz))
    pos = 0
    do dd = 1, Di
      mean = (Di - dd)**2
      do jj = lo(dd), hi(dd)
        pos = pos + 1
        r(pos) = int(dd - 1, int32)
        c(pos) = int(jj - 1, int32)
        v(pos) = fprob(real(jj - 1, real64), real(mean, real64))
      end do
    end do
    print '(a,i12,es25.16e3)', " Fsize", rank, real(nnz, real64)*8
  end subroutine

  ! Benchmark mode (argument 14 = 1, pqdts.f90:1581-1734): ten timed calls of
  ! each operator on a random row-stochastic X, printed as "t <op> seconds [value]".
  subroutine quick_benchmark()
    real(real64) :: t1, v1
    integer :: rep
    real(real64), allocatable :: G(:, :)
    allocate (G(M, N), diag(M), B(M, N))
    print '(a)', " quick benchmark:"
    call check(pqdts_put(ctx, SLOT_X, X, COL_MAJOR), "pqdts_put")
    do rep = 1, 10
      t1 = wtime()
      call check(pqdts_L(ctx, SLOT_X, v1), "pqdts_L")
      print '(a,2es25.16e3)', " t L", wtime() - t1, v1
    end do
    do rep = 1, 10
      t1 = wtime()
      call check(pqdts_dL(ctx, SLOT_X, SLOT_G), "pqdts_dL")
      ts = wtime() - t1
      call check(pqdts_get(ctx, SLOT_G, G, COL_MAJOR), "pqdts_get")
      print '(a,2es25.16e3)', " t dL", ts, G(1, 1)
    end do
    call time_op("pset", 1)
    call time_op("pcopy", 2)
    call time_op("pscale", 3)
    call time_op("pcopy_scale", 4)
    call time_op("paxpy", 5)
    call time_op("pgaxpy", 6)
    call time_op("pgaxpy2", 7)
    do rep = 1, 10
      t1 = wtime()
      call check(pqdts_pdot(ctx, SLOT_X, SLOT_G, v1), "pqdts_pdot")
      print '(a,2es25.16e3)', " t pdot", wtime() - t1, v1
    end do
    call time_op("doproj", 8)
    do rep = 1, 10
      t1 = wtime()
      call check(pqdts_Bdiag(ctx, diag), "pqdts_Bdiag")
      print '(a,2es25.16e3)', " t Bdiag", wtime() - t1, real(N, real64)*sum(diag(1:min(4_int64, M)))
    end do
    do rep = 1, 10
      t1 = wtime()
      call check(pqdts_BF(ctx, SLOT_G, SLOT_HP), "pqdts_BF")
      ts = wtime() - t1
      call check(pqdts_get(ctx, SLOT_HP, B, COL_MAJOR), "pqdts_get")
      print '(a,2es25.16e3)', " t BF", ts, sum(B(1:min(2_int64, M), :))
    end do
    deallocate (G, diag, B)
  end subroutine

  subroutine time_op(name, op)
    character(*), intent(in) :: name
    integer, intent(in) :: op
    real(real64) :: t1
    integer :: rep
    integer(c_int) :: rc
    do rep = 1, 10
      t1 = wtime()
      select case (op)
      case (1); rc = pqdts_pset(ctx, 1.0_c_double, SLOT_G)
      case (2); rc = pqdts_pcopy(ctx, SLOT_X, SLOT_G)
      case (3); rc = pqdts_pscale(ctx, -1.0_c_double, SLOT_G)
      case (4); rc = pqdts_pcopy_scale(ctx, -1.0_c_double, SLOT_X, SLOT_G)
      case (5); rc = pqdts_paxpy(ctx, 1.0_c_double, SLOT_X, SLOT_G)
      case (6); rc = pqdts_pgaxpy(ctx, 2.0_c_double, SLOT_X, 0.5_c_double, SLOT_G)
      case (7); rc = pqdts_pgaxpy2(ctx, 2.0_c_double, SLOT_X, 0.5_c_double, SLOT_G, SLOT_HP)
      case default; rc = pqdts_doproj(ctx, SLOT_G)
      end select
      call check(rc, name)
      print '(a,es25.16e3)', " t "//name, wtime() - t1
    end do
  end subroutine

  ! /proc/self/status copied to "status_<I12 rank> .out" (pqdts.f90:2208-2223);
  ! pqdts.py reads the VmHWM line from it (pqdts.py:85-88).
  subroutine write_status()
    character(len=256) :: line
    character(len=64) :: sname
    integer :: uin, uout, ios, p
    write (sname, '(a,i12,a)') "status_", rank, " .out"
    open (newunit=uout, file=trim(sname), status="replace", action="write")
    open (newunit=uin, file="/proc/self/status", status="old", action="read", iostat=ios)
    if (ios == 0) then
      do
        read (uin, '(a)', iostat=ios) line
        if (ios /= 0) exit
        p = index(line, ":")
        if (p > 0) then
          write (uout, '(1x,a,1x,a)') line(1:p), trim(adjustl(line(p + 1:)))
        end if
      end do
      close (uin)
    end if
    close (uout)
  end subroutine
end program
