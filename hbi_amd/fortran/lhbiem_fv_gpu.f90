! lhbiem_fv_gpu.f90 -- Fortran host of the fluid-driven cycle on the GPU:
! `lhbiem_fv_gpu <file.in>`, the counterpart of main_fv.f90.
!
! Reads main_fv's `key value` input (main_fv.f90:166-316, defaults 130-162),
! meshes a 2dp fault (coordinate2dp, 1327-1344) or a 2dph fault dipping at
! dipangle (coordinate2dph, 1346-1364), builds the lattice H-matrix through
! the HACApK call sites (m_HACApK_gpu; with the normal-stress matrix unless
! sigmaconst, 571-589), sets the initial condition of the setting
! (injectionq 1596-1643, injectionp 1548-1594, steadystate 1452-1498, thrust
! 1378-1450 with its models B and D, swarm 1645-1675), advances with hbi_fv_step (rkqs with the 4-component derivs,
! the implicit pore-pressure step, field update and stop controls on the
! device) and writes the reference's files: monitor<n>.dat (output_monitor,
! 1307-1311), time<n>.dat, event<n>.dat (1197-1239, decided from the
! per-step monitor values), the vel / slip / sigma / tau / ks / pf / qflow<n>.dat
! FP64 streams of every output step (1155-1198), xyz / ind<n>.dat
! (811-824) and the five local<n>-<id>.dat files (862-868, 1312-1316).  Output steps: every interval-th step exactly (a batch ends
! there) and the first step past each tout (hbi_fv_set_break ends the batch there); the step a stop
! control ends gets no rows (main_fv exits at 1136-1152, before its outputs).
! Keys beyond the reference's: vref (main_fv leaves it undeclared; 1e-6),
! leaf, precision, outbatch.
program lhbiem_fv_gpu
  use iso_c_binding
  use m_hbi_abi
  use m_HACApK_gpu
  implicit none
  real(8), parameter :: pi = 4 * datan(1.d0), rigid = 32.04d0, yr = 365 * 24 * 3600d0, &
    grav = real(9.80, 8)   ! the reference's g: a default-real literal
  character(len=256) :: input_file, param, pvalue, fname, problem, setting, bcl, bcr, evlaw, precision
  integer :: ios, number, ncellg, nstep, interval, leaf, outbatch, i, k, r, nb, nrow, kdone, u, nout
  real(8) :: ds0, tmax, tinj, kpmax, kpmin, eta, beta, phi0, q0, a0, b0, dc0, f00, tauinit, sigmainit
  real(8) :: velinit, pfinit, muinit, pbcl, pbcr, pinj, kL, kT, s0, eps_r, eps_h, vpl, dtout, velmin, velmax
  real(8) :: vref, Bs, dtinit, tout, meandisp0, moment, onset_time, kp0, lf, dipangle, qsum
  logical :: backslip, pressuredependent, pfconst, outpertime, slipping, outfield, stopped, sigmaconst
  character(len=256) :: model
  integer :: eventcount, hypoloc, lrtrn
  integer(c_int) :: rc
  type(st_HACApK_calc_entry) :: st_bemv
  type(st_HACApK_lcontrol) :: st_ctl
  type(st_HACApK_leafmtxp) :: st_leafmtxp_s, st_leafmtxp_n
  type(st_HACApK_LHp) :: st_LHp_s, st_LHp_n
  type(c_ptr) :: hn
  real(8), allocatable, target :: yel(:), yer(:), ang(:)
  real(8), allocatable :: qcum(:)
  type(st_HACApK_latticevec) :: st_sum
  real(8), allocatable :: coord(:, :), hist(:), y(:), pf(:), disp(:), vel(:), qflow(:)
  real(8), allocatable, target :: xcol(:), ycol(:), zcol(:), rake(:), xel(:), xer(:), ds(:)
  real(8), allocatable :: a(:), b(:), dc(:), f0(:), vplv(:), kpmaxv(:), kTv(:), phi(:), pfhyd(:), qin(:), area(:)
  real(8), allocatable :: sigma(:), sige(:), tau(:), ks(:), psi(:)
  type(hbi_fv_opts) :: fo
  type(hbi_fv_info) :: info
  type(c_ptr) :: fv
  integer(8) :: c0, c1, crate
  integer, parameter :: nl = 5
  integer :: locid(nl), nprobe, punit(nl)
  integer(c_int) :: probe_idx(nl)
  integer(c_long_long) :: nprow
  real(8), allocatable :: prow(:)

  call get_command_argument(1, input_file)
  open(33, file=input_file, status='old', iostat=ios)
  if (ios /= 0) then
    write(*, *) 'ERROR: Failed to open input file'
    stop 1
  end if
  ! defaults, main_fv.f90:130-162 (the keys it leaves unset start at 0 / F)
  number = 0; problem = ' '; ncellg = 0; ds0 = 0d0; nstep = 1000000; interval = 0
  setting = 'injectionq'; dtout = 1e3; eps_r = 1d-4; eps_h = 1d-4; velmax = 1d7; velmin = 1d-16; tmax = 1d12
  dtinit = 1d0; pressuredependent = .false.; pfconst = .false.; bcl = 'Neumann'; bcr = 'Neumann'
  evlaw = 'aging'; kT = 1e12; kL = 1e8; Bs = 0d0; outpertime = .false.; backslip = .false.
  tinj = 0d0; kpmax = 0d0; kpmin = 0d0; eta = 0d0; beta = 0d0; phi0 = 0d0; q0 = 0d0; a0 = 0d0; b0 = 0d0
  dc0 = 0d0; f00 = 0d0; tauinit = 0d0; sigmainit = 0d0; velinit = 0d0; pfinit = 0d0; muinit = 0d0
  pbcl = 0d0; pbcr = 0d0; pinj = 0d0; s0 = 0d0; vpl = 0d0; kp0 = 0d0
  vref = 1d-6; leaf = 32; precision = 'f64'; outbatch = 200
  model = ' '; dipangle = 0d0; sigmaconst = .false.
  do
    read(33, *, iostat=ios) param, pvalue
    if (ios /= 0) exit
    select case (trim(param))
    case ('problem'); read(pvalue, *) problem
    case ('ncellg'); read(pvalue, *) ncellg
    case ('nstep'); read(pvalue, *) nstep
    case ('filenumber'); read(pvalue, *) number
    case ('ds'); read(pvalue, *) ds0
    case ('velmax'); read(pvalue, *) velmax
    case ('velmin'); read(pvalue, *) velmin
    case ('a'); read(pvalue, *) a0
    case ('b'); read(pvalue, *) b0
    case ('dc'); read(pvalue, *) dc0
    case ('f0'); read(pvalue, *) f00
    case ('vpl'); read(pvalue, *) vpl
    case ('interval'); read(pvalue, *) interval
    case ('velinit'); read(pvalue, *) velinit
    case ('tauinit'); read(pvalue, *) tauinit
    case ('sigmainit'); read(pvalue, *) sigmainit
    case ('pfinit'); read(pvalue, *) pfinit
    case ('muinit'); read(pvalue, *) muinit
    case ('dtinit'); read(pvalue, *) dtinit
    case ('q0'); read(pvalue, *) q0
    case ('eta'); read(pvalue, *) eta
    case ('phi0'); read(pvalue, *) phi0
    case ('beta'); read(pvalue, *) beta
    case ('kpmax'); read(pvalue, *) kpmax
    case ('kpmin'); read(pvalue, *) kpmin
    case ('kp0'); read(pvalue, *) kp0
    case ('s0'); read(pvalue, *) s0
    case ('kL'); read(pvalue, *) kL
    case ('kT'); read(pvalue, *) kT
    case ('pbcl'); read(pvalue, *) pbcl
    case ('pbcr'); read(pvalue, *) pbcr
    case ('pinj'); read(pvalue, *) pinj
    case ('tmax'); read(pvalue, *) tmax
    case ('tinj'); read(pvalue, *) tinj
    case ('eps_r'); read(pvalue, *) eps_r
    case ('eps_h'); read(pvalue, *) eps_h
    case ('backslip'); read(pvalue, *) backslip
    case ('pressuredependent'); read(pvalue, *) pressuredependent
    case ('pfconst'); read(pvalue, *) pfconst
    case ('outpertime'); read(pvalue, *) outpertime
    case ('dtout'); read(pvalue, *) dtout
    case ('bcl'); read(pvalue, '(a)') bcl
    case ('bcr'); read(pvalue, '(a)') bcr
    case ('evlaw'); read(pvalue, '(a)') evlaw
    case ('setting'); read(pvalue, '(a)') setting
    case ('vref'); read(pvalue, *) vref
    case ('leaf'); read(pvalue, *) leaf
    case ('precision'); read(pvalue, '(a)') precision
    case ('outbatch'); read(pvalue, *) outbatch
    case ('model'); read(pvalue, '(a)') model
    case ('dipangle'); read(pvalue, *) dipangle
    case ('sigmaconst'); read(pvalue, *) sigmaconst
    case default
      write(*, *) 'WARNING: ', trim(param), ' is not read by this driver'
    end select
  end do
  close(33)
  if (trim(problem) /= '2dp' .and. trim(problem) /= '2dph') then
    write(*, *) 'ERROR: this driver meshes 2dp and 2dph faults'
    stop 1
  end if
  if (trim(problem) == '2dp') sigmaconst = .true.   ! main_fv.f90:573-576
  if (ncellg < 4 .or. ds0 <= 0d0) then
    write(*, *) 'ERROR: ncellg and ds must be set'
    stop 1
  end if

  ! ---------------------------------------------------------------- mesh --
  ! coordinate2dp (main_fv.f90:1327-1344) or coordinate2dph (1346-1364): a
  ! segment dipping at dipangle from the surface
  allocate(xcol(ncellg), ycol(ncellg), zcol(ncellg), rake(ncellg), xel(ncellg), xer(ncellg), ds(ncellg))
  allocate(yel(ncellg), yer(ncellg), ang(ncellg))
  ycol = 0d0; zcol = 0d0; rake = 0d0
  do i = 1, ncellg
    ds(i) = ds0
    if (trim(problem) == '2dph') then
      xel(i) = (i - 1) * ds0 * cos(dipangle * pi / 180)
      xer(i) = i * ds0 * cos(dipangle * pi / 180)
      yel(i) = (i - 1) * ds0 * sin(dipangle * pi / 180)
      yer(i) = i * ds0 * sin(dipangle * pi / 180)
      ycol(i) = 0.5d0 * (yel(i) + yer(i))
      ang(i) = datan2(yer(i) - yel(i), xer(i) - xel(i))
    else
      xel(i) = (i - 1) * ds0
      xer(i) = i * ds0
    end if
    xcol(i) = 0.5d0 * (xel(i) + xer(i))
  end do

  ! ------------------------------------------------------------ H-matrix --
  ! main_fv.f90:571-589: the shear matrix, and the normal-stress one unless
  ! sigma is constant (always for 2dp)
  lrtrn = HACApK_init(ncellg, st_ctl, st_bemv)
  st_ctl%param(21) = leaf
  select case (trim(precision))
  case ('f32'); st_ctl%param(P_PREC) = HBI_PREC_F32
  case ('mixed'); st_ctl%param(P_PREC) = HBI_PREC_MIXED
  case default; st_ctl%param(P_PREC) = HBI_PREC_F64
  end select
  st_bemv%problem = problem
  st_bemv%xcol => xcol; st_bemv%ycol => ycol; st_bemv%zcol => zcol; st_bemv%rake => rake
  st_bemv%xel => xel; st_bemv%xer => xer
  if (trim(problem) == '2dph') then
    st_bemv%yel => yel; st_bemv%yer => yer; st_bemv%ang => ang; st_bemv%ds => ds
  end if
  st_bemv%md = 's'; st_bemv%v = 's'
  allocate(coord(ncellg, 3))
  coord(:, 1) = xcol; coord(:, 2) = ycol; coord(:, 3) = zcol
  lrtrn = HACApK_generate(st_leafmtxp_s, st_bemv, st_ctl, coord, eps_h)
  if (lrtrn /= 0) stop 2
  lrtrn = HACApK_construct_LH(st_LHp_s, st_leafmtxp_s, st_bemv, st_ctl, coord, eps_h)
  lrtrn = HACApK_gen_lattice_vector(st_sum, st_leafmtxp_s, st_ctl)
  hn = c_null_ptr
  if (.not. sigmaconst) then
    st_bemv%v = 'n'
    lrtrn = HACApK_generate(st_leafmtxp_n, st_bemv, st_ctl, coord, eps_h)
    if (lrtrn /= 0) stop 2
    lrtrn = HACApK_construct_LH(st_LHp_n, st_leafmtxp_n, st_bemv, st_ctl, coord, eps_h)
    hn = st_LHp_n%h
  end if

  ! ------------------------------------------------- initial condition --
  allocate(a(ncellg), b(ncellg), dc(ncellg), f0(ncellg), vplv(ncellg), kpmaxv(ncellg), kTv(ncellg))
  allocate(phi(ncellg), pfhyd(ncellg), qin(ncellg), area(ncellg), sigma(ncellg), sige(ncellg), tau(ncellg))
  allocate(ks(ncellg), psi(ncellg), pf(ncellg), disp(ncellg), vel(ncellg), qflow(ncellg), y(4 * ncellg))
  a = a0; b = b0; dc = dc0; f0 = f00; vplv = vpl; area = ds0
  phi = phi0; pfhyd = 0d0; qin = 0d0; kTv = kT; kpmaxv = kpmax
  select case (trim(setting))
  case ('injectionq')   ! main_fv.f90:1596-1643
    qin(ncellg / 2) = q0
    ks = kpmin
    pf = pfinit
    sigma = sigmainit
    sige = sigma - pf
    tau = tauinit
    if (muinit /= 0d0) tau = sige * muinit
    fo%setting = 0
  case ('injectionp')   ! main_fv.f90:1548-1594
    ks = kpmin
    pf = pfinit
    pf(ncellg / 2) = pinj
    sigma = sigmainit
    sige = sigma - pf
    tau = sige * muinit
    fo%setting = 1
  case ('steadystate')  ! main_fv.f90:1452-1498: the linear pf profile carrying q0 along the fault
    lf = ncellg * ds0
    kp0 = kp0 * exp(sigmainit / s0)
    do i = 1, ncellg
      kTv(i) = kT
      if (abs(xcol(i) - lf / 2) > lf / 6) kTv(i) = kT * 100
    end do
    kpmaxv = kp0 * (1 + kL / kTv / vpl) - kpmin * kL / kTv / vpl
    ks = kp0
    pbcr = 0d0
    pbcl = pbcr + q0 * lf / (ks(1) * exp(-sigmainit / s0)) * eta * 1d-3
    pf = pbcl + (pbcr - pbcl) * xcol / lf
    sigma = sigmainit + pf
    sige = sigma - pf
    tau = sige * (f00 + (a - b) * log(velinit / vref)) * 1.001
    fo%setting = 2
  case ('thrust')       ! initcond_thrust, main_fv.f90:1378-1450 (its g is the default-real 9.80);
    ! indices as there: i is the element along the fault, qin(1950), qin(1600) and, for model D,
    ! qin(1200) are the sources; the library reads qin(2000) for the right boundary
    if (ncellg < 2000) then
      write(*, *) 'ERROR: setting thrust places sources at elements 1200-2000: ncellg >= 2000'
      stop 1
    end if
    do i = 1, ncellg
      a(i) = max(a0, a0 + (abs(i - 500) - 300) * 2d-5)
      if (trim(model) == 'B') a(i) = max(a0, a0 + (abs(i - 700) - 500) * 2d-5)
      a(i) = min(a(i), b0 + 0.002)
    end do
    qin(1950) = q0
    if (trim(model) == 'D') qin(1200) = -19 * q0
    qin(1600) = 19 * q0
    allocate(qcum(ncellg))
    qsum = sum(qin)
    do i = 1, ncellg
      qcum(i) = qsum - qin(i)
      qsum = qcum(i)
      kTv(i) = min(1e0 * exp(1d4 / (300d0 + 12d0 * ycol(i))), 1d12)
      ks(i) = (kpmax * vpl / kL + kpmin / kTv(i)) / (vpl / kL + 1 / kTv(i))
    end do
    sigma = grav * 2.6d3 * (ycol + 0.2) * 1d-3
    pfhyd = grav * 1.0d3 * (ycol + 0.2) * 1d-3
    pf(1) = pfhyd(1)
    sige(1) = sigma(1) - pf(1)
    pbcl = pfhyd(1)
    do i = 2, ncellg
      sige(i) = sige(i - 1) + ds0 * (sin(dipangle * pi / 180d0) * (2.6d3 - 1.0d3) * grav * 1d-3 &
                - qcum(i) * eta / ks(i) * 1d-3 * exp(sige(i - 1) / s0))
      pf(i) = sigma(i) - sige(i)
    end do
    pbcr = pf(ncellg)
    tau = sige * (f0 + (a - b) * log(velinit / vref)) * 1.001
    fo%setting = 3
  case ('swarm')        ! main_fv.f90:1645-1675
    kTv = 1e12
    ks = kpmin
    ks(1:min(10, ncellg)) = kpmax
    pf = pfinit
    pf(1:min(10, ncellg)) = pbcl
    sigma = sigmainit
    sige = sigma - pf
    tau = sige * muinit
    fo%setting = 4
  case default
    write(*, *) 'ERROR: setting ', trim(setting), ' is not built in this driver'
    stop 1
  end select
  vel = velinit
  psi = a * dlog(2 * vref / vel * sinh(tau / sige / a))
  disp = 0d0
  qflow = q0
  do i = 1, ncellg
    y(4 * i - 3) = psi(i); y(4 * i - 2) = tau(i); y(4 * i - 1) = sige(i); y(4 * i) = ks(i)
  end do
  fo%law = merge(1, 0, trim(evlaw) == 'slip')
  fo%backslip = merge(1, 0, backslip)
  fo%prec = st_ctl%param(P_PREC)
  fo%pressure_dependent = merge(1, 0, pressuredependent)
  fo%pfconst = merge(1, 0, pfconst)
  fo%bcl = merge(1, 0, trim(bcl) == 'Dirichlet')
  fo%bcr = merge(1, 0, trim(bcr) == 'Dirichlet')
  fo%outpertime = merge(1, 0, outpertime)
  fo%vref = vref; fo%eps_r = eps_r; fo%hdiv = 0d0; fo%ds0 = ds0; fo%kL = kL; fo%kpmin = kpmin; fo%s0 = s0
  fo%eta = eta; fo%beta = beta; fo%Bs = Bs; fo%pbcl = pbcl; fo%pbcr = pbcr; fo%q0 = q0; fo%pinj = pinj
  fo%tinj_s = tinj * 365 * 24 * 3600; fo%dtout_s = dtout * 365 * 24 * 3600
  fo%velmin = velmin; fo%velmax = velmax; fo%tmax_s = tmax * 365 * 24 * 3600
  ! the state arrays go in cluster order (np = 1: lodc is the cluster permutation)
  rc = hbi_fv_create(st_LHp_s%h, hn, fo, perm(a), perm(b), perm(dc), perm(f0), perm(vplv), &
                     perm(kpmaxv), perm(kTv), perm(phi), perm(pfhyd), perm(qin), perm(area), fv)
  if (rc /= 0) then
    write(*, *) 'ERROR: ', trim(hbi_error_text())
    stop 2
  end if
  call permute_state()
  rc = hbi_fv_set_state(fv, y, pf, disp, 0d0, dtinit)

  ! ------------------------------------------------------------- outputs --
  call execute_command_line('mkdir -p output')
  write(fname, '("output/ind",i0,".dat")') number
  open(newunit=u, file=fname, form='unformatted', access='stream', status='replace')
  write(u) st_sum%lodc(1:ncellg)
  close(u)
  write(fname, '("output/xyz",i0,".dat")') number
  open(newunit=u, file=fname, status='replace')
  do i = 1, ncellg
    write(u, '(3e15.6)') xcol(st_sum%lodc(i)), ycol(st_sum%lodc(i)), zcol(st_sum%lodc(i))
  end do
  close(u)
  call open_stream(81, 'vel'); call open_stream(82, 'slip'); call open_stream(83, 'sigma')
  call open_stream(84, 'tau'); call open_stream(85, 'ks'); call open_stream(86, 'pf')
  call open_stream(87, 'qflow')
  write(fname, '("output/time",i0,".dat")') number
  open(50, file=fname, status='replace')
  write(fname, '("output/event",i0,".dat")') number
  open(51, file=fname, status='replace')
  write(fname, '("output/monitor",i0,".dat")') number
  open(52, file=fname, status='replace')
  ! the initial condition (main_fv.f90:950-971): fields, k = 0 rows
  write(81) vel_of_state()
  write(82) disp
  write(83) y(3::4)
  write(84) y(2::4)
  write(85) kp_of_state()
  write(86) pf
  write(87) qflow
  write(52, '(i7,f19.4,8e16.5,f16.4,i6)') 0, 0d0, log10(maxval(abs(vel_of_state()))), 0d0, &
    sum(y(2::4) / y(3::4) * ds0) / (ds0 * ncellg), maxval(y(3::4)), minval(y(3::4)), &
    xcol(maxloc(abs(vel_of_state()), 1)), 0d0, 0d0, 0d0, 0
  write(50, '(i7,f19.4)') 0, 0d0
  ! local output points, main_fv.f90:862-868: cluster-order indices
  nprobe = 0
  do i = 1, nl
    locid(i) = int(dble(i - 0.5) / nl * ncellg)
    if (locid(i) < 1 .or. locid(i) > ncellg) cycle
    nprobe = nprobe + 1
    probe_idx(nprobe) = locid(i) - 1
    punit(nprobe) = 52 + i
    write(fname, '("output/local",i0,"-",i0,".dat")') number, locid(i)
    open(52 + i, file=fname, status='replace')
  end do
  rc = hbi_fv_set_probes(fv, int(nprobe, c_int), probe_idx)
  allocate(prow(9 * max(nprobe, 1) * outbatch))

  allocate(hist(10 * outbatch))
  tout = dtout * 365 * 24 * 60 * 60
  slipping = .false.; eventcount = 0; meandisp0 = 0d0; onset_time = 0d0; hypoloc = 0
  kdone = 0
  call system_clock(c0, crate)
  do while (kdone < nstep)
    nb = min(outbatch, nstep - kdone)
    if (interval > 0) nb = min(nb, interval - mod(kdone, interval))
    ! the call also ends at the first step past tout (field output, main_fv.f90:1160-1163)
    rc = hbi_fv_set_break(fv, tout, huge(1d0), 0d0)
    rc = hbi_fv_step(fv, int(nb, c_int), hist, info)
    call system_clock(c1)
    nrow = int(info%accepted) - kdone
    ! a stop control ends main_fv's loop before the outputs of that step
    ! (main_fv.f90:1136-1152 come before 1156-1239): no rows for it
    stopped = info%status == 1 .or. info%status == 3 .or. info%status == 4
    nout = nrow
    if (stopped) nout = max(nrow - 1, 0)
    do r = 1, nout
      k = kdone + r
      ! output_monitor, main_fv.f90:1307-1311 (xcol indexed by the cluster
      ! index of max |V|, as there)
      write(52, '(i7,f19.4,8e16.5,f16.4,i6)') k, hist(10 * r - 9), hist(10 * r - 8), hist(10 * r - 7), &
        hist(10 * r - 6), hist(10 * r - 5), hist(10 * r - 4), xcol(int(hist(10 * r - 3))), hist(10 * r - 2), &
        hist(10 * r - 1), real(c1 - c0, 8) / real(crate, 8), int(hist(10 * r))
      ! event list, main_fv.f90:1197-1239
      if (.not. slipping .and. 10d0 ** hist(10 * r - 8) > 1d-2) then
        slipping = .true.
        eventcount = eventcount + 1
        meandisp0 = hist(10 * r - 7)
        hypoloc = int(hist(10 * r - 3))
        onset_time = hist(10 * r - 9)
      end if
      if (slipping .and. 10d0 ** hist(10 * r - 8) < 1d-3) then
        slipping = .false.
        moment = hist(10 * r - 7) - meandisp0
        write(51, '(i0,i0,f17.2,i5,e15.6,f14.4)') eventcount, k, onset_time, hypoloc, moment, &
          (log10(moment * rigid * ds0 * ncellg) + 5.9) / 1.5
      end if
    end do
    if (nprobe > 0) call write_local()
    kdone = int(info%accepted)
    outfield = interval > 0 .and. mod(kdone, max(interval, 1)) == 0 .and. nrow > 0
    if (info%t > tout .and. .not. stopped) then
      outfield = .true.
      tout = info%t + dtout * 365 * 24 * 60 * 60
    end if
    if (outfield .and. .not. stopped) call write_fields()
    if (info%status /= 0) then
      select case (info%status)
      case (1); write(*, *) 'slip rate outside [velmin, velmax] at time step ', kdone
      case (3); write(*, *) 'time exceeds tmax'
      case (4); write(*, *) 'normal stress below 0.2MPa at time step ', kdone
      case default; write(*, *) 'ERROR: ', trim(hbi_error_text())
      end select
      exit
    end if
    if (nrow == 0) exit
  end do
  call system_clock(c1)
  write(*, *) 'time(s)', real(c1 - c0, 8) / real(crate, 8), ' steps', kdone, ' t_end', info%t
  do i = 81, 87
    close(i)
  end do
  close(50); close(51); close(52)
  do i = 1, nprobe
    close(punit(i))
  end do
  rc = hbi_fv_destroy(fv)
  lrtrn = HACApK_free_leafmtxp(st_leafmtxp_s)
  if (.not. sigmaconst) lrtrn = HACApK_free_leafmtxp(st_leafmtxp_n)
  lrtrn = HACApK_finalize(st_ctl)

contains

  ! output_local rows (main_fv.f90:1312-1316) of the last batch: k, t,
  ! log10 V, disp, sigma_e, tau, pf, sigma, mu, psi, kp
  subroutine write_local()
    integer :: j, rr, o
    rc = hbi_fv_get_probes(fv, nprow, prow)
    do rr = 1, min(int(nprow), nout)
      do j = 1, nprobe
        o = 9 * (nprobe * (rr - 1) + j - 1)
        write(punit(j), '(i7,f19.4,9e16.6)') kdone + rr, hist(10 * rr - 9), log10(prow(o + 1)), prow(o + 2), &
          prow(o + 3), prow(o + 4), prow(o + 5), prow(o + 6), prow(o + 7), prow(o + 8), prow(o + 9)
      end do
    end do
  end subroutine write_local

  ! a mesh-order array in cluster order
  function perm(v) result(w)
    real(8), intent(in) :: v(:)
    real(8) :: w(size(v))
    integer :: j
    do j = 1, size(v)
      w(j) = v(st_sum%lodc(j))
    end do
  end function perm

  subroutine permute_state()
    real(8), allocatable :: y0(:)
    integer :: j, jj
    y0 = y
    do j = 1, ncellg
      jj = st_sum%lodc(j)
      y(4 * j - 3:4 * j) = y0(4 * jj - 3:4 * jj)
    end do
    pf = perm(pf)
    disp = perm(disp)
    qflow = perm(qflow)
  end subroutine permute_state

  ! V of the state y (cluster order), main_fv.f90:1066
  function vel_of_state() result(v)
    real(8) :: v(ncellg)
    integer :: j
    do j = 1, ncellg
      v(j) = 2 * vref * exp(-y(4 * j - 3) / a(st_sum%lodc(j))) * sinh(y(4 * j - 2) / y(4 * j - 1) / a(st_sum%lodc(j)))
    end do
  end function vel_of_state

  ! the fields of an output step (main_fv.f90:1175-1190) and its time row
  subroutine write_fields()
    rc = hbi_fv_get_state(fv, y, pf, disp, vel, qflow, info)
    write(50, '(i7,f19.4)') kdone, info%t
    write(81) vel
    write(82) disp
    write(83) y(3::4)
    write(84) y(2::4)
    write(85) kp_of_state()
    write(86) pf
    write(87) qflow
  end subroutine write_fields

  ! kp: ks, or ks exp(-sigma_e / s0) with pressure-dependent permeability (main_fv.f90:1060-1068)
  function kp_of_state() result(kp)
    real(8) :: kp(ncellg)
    if (pressuredependent) then
      kp = y(4::4) * exp(-y(3::4) / s0)
    else
      kp = y(4::4)
    end if
  end function kp_of_state

  subroutine open_stream(un, name)
    integer, intent(in) :: un
    character(len=*), intent(in) :: name
    character(len=256) :: path
    write(path, '("output/",a,i0,".dat")') name, number
    open(un, file=path, form='unformatted', access='stream', status='replace')
  end subroutine open_stream

end program lhbiem_fv_gpu
