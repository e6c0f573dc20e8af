! lhbiem_fv3dp_gpu.f90 -- Fortran host of the fluid-driven cycle with the
! pore pressure on the fault's own 2-D grid: `lhbiem_fv3dp_gpu <file.in>`,
! the counterpart of main_fv3dp.f90.
!
! Reads main_fv3dp's `key value` input (main_fv3dp.f90:159-317, defaults
! 120-155; `bc` / `pbc` set all four sides, 319-323), meshes the imax x jmax
! squares of a 3dp fault (coordinate3dp, 1174-1202), builds the shear
! H-matrix (3dp: sigma constant, 533-541), takes friction from the keys or
! parameter_file (549-560), sets the initial condition of the setting
! (injection 1020-1061, ss 1063-1099, thrust 979-1018) and the wells of
! injection_file (input_well, 1101-1118), advances with hbi_fv_step in grid
! mode (rkqs over the 4-component derivs with hnext <= dtmax, the implicit
! pore-pressure step on the 2-D grid, the minsig rule and the stop controls
! on the device) and writes the reference's files, in mesh order as
! main_fv3dp keeps its arrays: monitor<n>.dat (output_monitor, 951-955),
! time<n>.dat, event<n>.dat and EQslip<n>.dat (871-905), the vel / slip /
! sigma / tau / ks / pf / qflow<n>.dat FP64 streams of the initial state and
! of every output step (676-685, 820-861), local<n>-<i>.dat at the five points
! of find_locid (642-646, output_local 956-960).
! Output steps: every interval-th step and the first step past each tout (a
! step call ends there through hbi_fv_set_break).  An event's onset (max V
! > 1e-2) and end (< 1e-3) end a step call as well, so its coseismic slip is
! the difference of the displacement fields of those two steps.  The step a
! stop control ends gets no rows (main_fv3dp exits at 801-818, before its
! outputs).  Keys beyond the reference's: vref (main_fv3dp fixes 1e-6), leaf,
! precision, outbatch.  interval 0 (unset) gives no interval outputs.
program lhbiem_fv3dp_gpu
  use iso_c_binding
  use m_hbi_abi
  use m_HACApK_gpu
  implicit none
  real(8), parameter :: pi = 4 * datan(1.d0), rigid = 32.04d0, yr = 365 * 24 * 3600d0, &
    grav = real(9.80, 8)   ! the reference's g: a default-real literal
  character(len=256) :: input_file, param, pvalue, fname, problem, setting, bc, bcl, bcr, bct, bcb, evlaw
  character(len=256) :: precision, parameter_file, injection_file
  integer :: ios, number, imax, jmax, ncellg, nstep, interval, leaf, outbatch, i, j, k, r, nb, nrow, kdone
  integer :: nout, hypoloc, eventcount, lrtrn, nwell, w, l
  real(8) :: ds0, tmax, dtmax, kpmax, kpmin, kp0, eta, beta, phi0, q0, a0, b0, dc0, f00, tauinit, sigmainit
  real(8) :: velinit, pfinit, muinit, pbc, pbcl, pbcr, kL, kT, s0, eps_r, eps_h, vpl, dtout, velmin, velmax
  real(8) :: vref, dtinit, tout, meandisp0, moment, onset_time, minsig, dipangle, lf, sg, vnow
  logical :: backslip, pressuredependent, parameterfromfile, injectionfromfile, slipping, outfield, stopped
  integer(c_int) :: rc
  type(st_HACApK_calc_entry) :: st_bemv
  type(st_HACApK_lcontrol) :: st_ctl
  type(st_HACApK_leafmtxp) :: st_leafmtxp_s
  type(st_HACApK_LHp) :: st_LHp_s
  type(st_HACApK_latticevec) :: st_sum
  real(8), allocatable :: coord(:, :), ds(:), hist(:), y(:), pf(:), disp(:), vel(:), qflow(:), idisp(:)
  real(8), allocatable, target :: xcol(:), ycol(:), zcol(:), xq(:, :)
  real(8), allocatable :: a(:), b(:), dc(:), f0(:), vplv(:), kpmaxv(:), kTv(:), phi(:), pfhyd(:), qin(:), area(:)
  real(8), allocatable :: sigma(:), sige(:), tau(:), ks(:), psi(:), mu(:), sigmaed(:)
  integer(c_int), allocatable :: iwell(:), jwell(:), kleng(:), inv(:)
  real(8), allocatable :: qvals(:), qtimes(:)
  type(hbi_fv_opts) :: fo
  type(hbi_fv_info) :: info
  type(c_ptr) :: fv
  integer(8) :: c0, c1, crate
  integer, parameter :: nl = 5
  integer :: locid(nl), punit(nl)
  integer(c_int) :: probe_idx(nl)
  integer(c_long_long) :: nprow
  real(8), allocatable :: prow(:)

  call get_command_argument(1, input_file)
  open(33, file=input_file, status='old', iostat=ios)
  if (ios /= 0) then
    write(*, *) 'Failed to open inputfile'
    stop 1
  end if
  ! defaults, main_fv3dp.f90:120-155 (the keys it leaves unset start at 0 / F)
  number = 0; problem = ' '; imax = 0; jmax = 0; nstep = 1000000; interval = 0; ds0 = 0d0
  setting = 'injection'; eps_r = 1d-4; eps_h = 1d-4; vref = 1d-6; velmax = 1d7; velmin = 1d-16; tmax = 1d12
  dipangle = 0d0; minsig = 0d0; dtinit = 1d0; muinit = 0d0; dtmax = 1d10; dtout = 0d0
  parameterfromfile = .false.; injectionfromfile = .false.; pressuredependent = .false.
  bc = 'default'; bcb = 'Neumann'; bct = 'Neumann'; bcl = 'Neumann'; bcr = 'Neumann'; evlaw = 'aging'
  backslip = .false.; kpmax = 0d0; kpmin = 0d0; kp0 = 0d0; eta = 0d0; beta = 0d0; phi0 = 0d0; q0 = 0d0
  a0 = 0d0; b0 = 0d0; dc0 = 0d0; f00 = 0d0; tauinit = 0d0; sigmainit = 0d0; velinit = 0d0; pfinit = 0d0
  pbc = 0d0; pbcl = 0d0; pbcr = 0d0; kL = 0d0; kT = 0d0; s0 = 0d0; vpl = 0d0
  parameter_file = ' '; injection_file = ' '
  leaf = 32; precision = 'f64'; outbatch = 200
  do
    read(33, *, iostat=ios) param, pvalue
    if (ios /= 0) exit
    select case (trim(param))
    case ('problem'); read(pvalue, *) problem
    case ('imax'); read(pvalue, *) imax
    case ('jmax'); read(pvalue, *) jmax
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
    case ('pbc'); read(pvalue, *) pbc
    case ('pbcl'); read(pvalue, *) pbcl
    case ('pbcr'); read(pvalue, *) pbcr
    case ('tmax'); read(pvalue, *) tmax
    case ('dtmax'); read(pvalue, *) dtmax
    case ('eps_r'); read(pvalue, *) eps_r
    case ('eps_h'); read(pvalue, *) eps_h
    case ('backslip'); read(pvalue, *) backslip
    case ('pressuredependent'); read(pvalue, *) pressuredependent
    case ('minsig'); read(pvalue, *) minsig
    case ('dipangle'); read(pvalue, *) dipangle
    case ('dtout'); read(pvalue, *) dtout
    case ('parameterfromfile'); read(pvalue, *) parameterfromfile
    case ('parameter_file'); read(pvalue, '(a)') parameter_file
    case ('injectionfromfile'); read(pvalue, *) injectionfromfile
    case ('injection_file'); read(pvalue, '(a)') injection_file
    case ('bc'); read(pvalue, '(a)') bc
    case ('bcl'); read(pvalue, '(a)') bcl
    case ('bcr'); read(pvalue, '(a)') bcr
    case ('bct'); read(pvalue, '(a)') bct
    case ('bcb'); read(pvalue, '(a)') bcb
    case ('evlaw'); read(pvalue, '(a)') evlaw
    case ('setting'); read(pvalue, '(a)') setting
    case ('vref'); read(pvalue, *) vref
    case ('leaf'); read(pvalue, *) leaf
    case ('precision'); read(pvalue, '(a)') precision
    case ('outbatch'); read(pvalue, *) outbatch
    ! read by main_fv3dp and not used by its 3dp cycle (pfconst: the test
    ! around the pressure step is commented out, 753-759)
    case ('ncellg', 'loc', 'a1', 'load', 'sr', 'geometry_file', 'ksinit', 'sparam', 'cd', 'td', 'slipevery', &
          'limitsigma', 'sigmaconst', 'foward', 'inverse', 'geofromfile', 'maxsig', 'crake', 'outpertime', &
          'restart', 'latticeh', 'pfconst')
    case default
      write(*, *) 'WARNING: ', trim(param), ' is an unknown parameter'
    end select
  end do
  close(33)
  if (trim(bc) /= 'default') then
    bcl = bc; bcr = bc; bct = bc; bcb = bc
    pbcl = pbc; pbcr = pbc
  end if
  if (trim(problem) /= '3dp') then
    write(*, *) 'ERROR: this driver meshes 3dp faults (main_fv3dp gives 3dph no dip angles in its kernel set-up)'
    stop 1
  end if
  if (imax < 4 .or. jmax < 4 .or. ds0 <= 0d0) then
    write(*, *) 'ERROR: imax, jmax >= 4 and ds must be set'
    stop 1
  end if
  ncellg = imax * jmax

  ! ---------------------------------------------------------------- mesh --
  ! coordinate3dp, main_fv3dp.f90:1174-1202: element k = (i - 1) jmax + j
  allocate(xcol(ncellg), ycol(ncellg), zcol(ncellg), ds(ncellg), xq(ncellg, 8))
  ycol = 0d0
  do i = 1, imax
    do j = 1, jmax
      k = (i - 1) * jmax + j
      xcol(k) = (i - imax / 2 - 0.5d0) * ds0
      zcol(k) = -(j - jmax / 2 - 0.5d0) * ds0
      xq(k, 1) = xcol(k) + 0.5d0 * ds0; xq(k, 2) = xcol(k) - 0.5d0 * ds0
      xq(k, 3) = xcol(k) - 0.5d0 * ds0; xq(k, 4) = xcol(k) + 0.5d0 * ds0
      xq(k, 5) = zcol(k) + 0.5d0 * ds0; xq(k, 6) = zcol(k) + 0.5d0 * ds0
      xq(k, 7) = zcol(k) - 0.5d0 * ds0; xq(k, 8) = zcol(k) - 0.5d0 * ds0
    end do
  end do
  ds = ds0 * ds0

  ! ------------------------------------------------------------ H-matrix --
  lrtrn = HACApK_init(ncellg, st_ctl, st_bemv)
  st_ctl%param(21) = leaf
  select case (trim(precision))
  case ('f32'); st_ctl%param(P_PREC) = HBI_PREC_F32
  case ('mixed'); st_ctl%param(P_PREC) = HBI_PREC_MIXED
  case default; st_ctl%param(P_PREC) = HBI_PREC_F64
  end select
  st_bemv%problem = problem
  st_bemv%xcol => xcol; st_bemv%ycol => ycol; st_bemv%zcol => zcol
  st_bemv%xs1 => xq(:, 1); st_bemv%xs2 => xq(:, 2); st_bemv%xs3 => xq(:, 3); st_bemv%xs4 => xq(:, 4)
  st_bemv%zs1 => xq(:, 5); st_bemv%zs2 => xq(:, 6); st_bemv%zs3 => xq(:, 7); st_bemv%zs4 => xq(:, 8)
  st_bemv%md = 's'; st_bemv%v = 's'
  allocate(coord(ncellg, 3))
  coord(:, 1) = xcol; coord(:, 2) = ycol; coord(:, 3) = zcol
  lrtrn = HACApK_generate(st_leafmtxp_s, st_bemv, st_ctl, coord, eps_h)
  if (lrtrn /= 0) stop 2
  lrtrn = HACApK_construct_LH(st_LHp_s, st_leafmtxp_s, st_bemv, st_ctl, coord, eps_h)
  lrtrn = HACApK_gen_lattice_vector(st_sum, st_leafmtxp_s, st_ctl)
  allocate(inv(ncellg))
  do j = 1, ncellg
    inv(st_sum%lodc(j)) = j
  end do

  ! ------------------------------------------------- initial condition --
  allocate(a(ncellg), b(ncellg), dc(ncellg), f0(ncellg), vplv(ncellg), kpmaxv(ncellg), kTv(ncellg))
  allocate(phi(ncellg), pfhyd(ncellg), qin(ncellg), area(ncellg), sigma(ncellg), sige(ncellg), tau(ncellg))
  allocate(ks(ncellg), psi(ncellg), pf(ncellg), disp(ncellg), vel(ncellg), qflow(ncellg), y(4 * ncellg))
  allocate(mu(ncellg), idisp(ncellg))
  a = a0; b = b0; dc = dc0; f0 = f00
  if (parameterfromfile) then   ! main_fv3dp.f90:554-560, mesh order
    open(99, file=parameter_file, status='old', iostat=ios)
    if (ios /= 0) then
      write(*, *) 'ERROR: Failed to open parameter file ', trim(parameter_file)
      stop 1
    end if
    do i = 1, ncellg
      read(99, *) a(i), b(i), dc(i), f0(i)
    end do
    close(99)
  end if
  ! plain means in the monitor (main_fv3dp.f90:795-799): unit weights
  vplv = vpl; area = 1d0; qin = 0d0; pfhyd = 0d0; mu = 0d0
  kpmaxv = kpmax
  select case (trim(setting))
  case ('injection')   ! initcond_injection, main_fv3dp.f90:1020-1061 (it leaves mu unset)
    phi = phi0
    ks = kpmin
    kTv = kT
    sigma = sigmainit
    pf = 0d0
    sige = sigma - pf
    vel = velinit
    tau = tauinit
    if (muinit /= 0d0) tau = sige * muinit
    fo%setting = HBI_FV_INJECTIONQ
  case ('ss')          ! initcond_bgflow, main_fv3dp.f90:1063-1099
    lf = imax * ds0
    kTv = kT
    kpmaxv = kp0 * (1 + kL / kT / vpl) - kpmin * kL / kT / vpl
    ks = kp0
    pbcr = 0d0
    pbcl = pbcr + q0 * lf / ks(1) * eta * 1d-3
    phi = phi0
    pf = pbcl + (pbcr - pbcl) * xcol / lf
    sigma = sigmainit + pf
    sige = sigma - pf
    vel = velinit
    tau = sige * (f0 + (a - b) * log(vel / vref)) * 1.001
    mu = tau / sige
    fo%setting = HBI_FV_STEADYSTATE
  case ('thrust')      ! initcond_thrust, main_fv3dp.f90:979-1018 (its g is the default-real 9.80)
    sigma = grav * 2.6d3 * (zcol + 0.2) * 1d-3
    pfhyd = grav * 1.0d3 * (zcol + 0.2) * 1d-3
    allocate(sigmaed(jmax))
    sigmaed(1) = sigma(1) - pfhyd(1)
    pbcl = pfhyd(1)
    kTv = kT
    kpmaxv = kp0 * (1 + kL / kT / vpl) - kpmin * kL / kT / vpl
    ks = kp0
    do j = 2, jmax
      sigmaed(j) = sigmaed(j - 1) + ds0 * (sin(dipangle * pi / 180d0) * (2.6d3 - 1.0d3) * grav * 1d-3 &
                   - q0 * eta / kp0 * 1d-3 * exp(sigmaed(j - 1) / s0))
    end do
    do l = 1, ncellg
      i = (l - 1) / jmax + 1
      j = l - (i - 1) * jmax
      sige(l) = sigmaed(j)
      pf(l) = sigma(l) - sige(l)
      phi(l) = phi0
    end do
    pbcr = pf(ncellg)
    vel = velinit
    tau = sige * (f0 + (a - b) * log(vel / vref)) * 1.001
    mu = tau / sige
    fo%setting = HBI_FV_THRUST
  case default
    write(*, *) 'ERROR: setting ', trim(setting), ' is not one of main_fv3dp''s (injection, ss, thrust)'
    stop 1
  end select
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
  fo%pfconst = 0
  fo%bcl = merge(1, 0, trim(bcl) == 'Dirichlet')
  fo%bcr = merge(1, 0, trim(bcr) == 'Dirichlet')
  fo%bct = merge(1, 0, trim(bct) == 'Dirichlet')
  fo%bcb = merge(1, 0, trim(bcb) == 'Dirichlet')
  fo%outpertime = 0
  fo%vref = vref; fo%eps_r = eps_r; fo%hdiv = 0d0; fo%ds0 = ds0; fo%kL = kL; fo%kpmin = kpmin; fo%s0 = s0
  fo%eta = eta; fo%beta = beta; fo%Bs = 0d0; fo%pbcl = pbcl; fo%pbcr = pbcr; fo%q0 = q0; fo%pinj = 0d0
  fo%tinj_s = 0d0; fo%dtout_s = 0d0
  fo%velmin = velmin; fo%velmax = velmax; fo%tmax_s = tmax * yr
  fo%imax = imax; fo%jmax = jmax; fo%dtmax = dtmax; fo%minsig = minsig; fo%phi0 = phi0
  rc = hbi_fv_create(st_LHp_s%h, c_null_ptr, fo, perm(a), perm(b), perm(dc), perm(f0), perm(vplv), &
                     perm(kpmaxv), perm(kTv), perm(phi), perm(pfhyd), perm(qin), perm(area), fv)
  if (rc /= 0) then
    write(*, *) 'ERROR: ', trim(hbi_error_text())
    stop 2
  end if
  call check(hbi_fv_set_state(fv, permute_state(), perm(pf), perm(disp), 0d0, dtinit), 'hbi_fv_set_state')
  if (injectionfromfile) then   ! input_well, main_fv3dp.f90:1101-1118
    open(77, file=injection_file, status='old', iostat=ios)
    if (ios /= 0) then
      write(*, *) 'ERROR: Failed to open injection file ', trim(injection_file)
      stop 1
    end if
    read(77, *) nwell
    write(*, *) 'nwell', nwell
    allocate(iwell(nwell), jwell(nwell), kleng(nwell), qvals(50 * nwell), qtimes(50 * nwell))
    qvals = 0d0; qtimes = 0d0
    do w = 1, nwell
      read(77, *) iwell(w), jwell(w), kleng(w)
      read(77, *) qvals(50 * (w - 1) + 1:50 * (w - 1) + kleng(w))
      read(77, *) qtimes(50 * (w - 1) + 1:50 * (w - 1) + kleng(w))
    end do
    close(77)
    call check(hbi_fv_set_wells(fv, int(nwell, c_int), iwell, jwell, kleng, qvals, qtimes), 'hbi_fv_set_wells')
  end if

  ! ------------------------------------------------------------- outputs --
  call execute_command_line('mkdir -p output')
  call open_stream(81, 'vel'); call open_stream(82, 'slip'); call open_stream(83, 'sigma')
  call open_stream(84, 'tau'); call open_stream(85, 'ks'); call open_stream(86, 'pf')
  call open_stream(87, 'qflow'); call open_stream(88, 'EQslip')
  write(fname, '("output/time",i0,".dat")') number
  open(50, file=fname, status='replace')
  write(fname, '("output/event",i0,".dat")') number
  open(51, file=fname, status='replace')
  write(fname, '("output/monitor",i0,".dat")') number
  open(52, file=fname, status='replace')
  ! find_locid, main_fv3dp.f90:642-646, 969-975: mesh elements, probed by
  ! their cluster-order index
  locid(1) = (imax / 2 - 1) * jmax + jmax / 2
  locid(2) = (imax / 4 - 1) * jmax + jmax / 2
  locid(3) = (imax * 3 / 4 - 1) * jmax + jmax / 2
  locid(4) = (imax / 2 - 1) * jmax + jmax / 4
  locid(5) = (imax / 2 - 1) * jmax + jmax * 3 / 4
  do i = 1, nl
    probe_idx(i) = inv(locid(i)) - 1
    punit(i) = 52 + i
    write(fname, '("output/local",i0,"-",i0,".dat")') number, i
    open(52 + i, file=fname, status='replace')
  end do
  call check(hbi_fv_set_probes(fv, int(nl, c_int), probe_idx), 'hbi_fv_set_probes')
  allocate(prow(9 * nl * outbatch))
  ! the initial rows (main_fv3dp.f90:665-685; mu as the setting leaves it)
  sg = ds0 * ds0 * ncellg
  call system_clock(c0, crate)
  write(50, '(i7,f19.4)') 0, 0d0
  write(52, '(i7,f19.4,7e16.5,f16.4,i6)') 0, 0d0, log10(maxval(vel)), 0d0, sum(mu) / ncellg, maxval(sige), &
    minval(sige), 0d0, 0d0, 0d0, 0
  write(81) vel
  write(82) disp
  write(83) sige
  write(84) tau
  if (pressuredependent) then
    write(85) ks * exp(-sige / s0)
  else
    write(85) ks
  end if
  write(86) pf
  write(87) qflow

  allocate(hist(10 * outbatch))
  tout = dtout * yr
  slipping = .false.; eventcount = 0; meandisp0 = 0d0; onset_time = 0d0; hypoloc = 0
  kdone = 0
  do while (kdone < nstep)
    nb = min(outbatch, nstep - kdone)
    if (interval > 0) nb = min(nb, interval - mod(kdone, interval))
    ! the call also ends at the first step past tout and at an event's onset
    ! or end (main_fv3dp.f90:829-832, 871-905)
    if (slipping) then
      call check(hbi_fv_set_break(fv, tout, huge(1d0), 1d-3), 'hbi_fv_set_break')
    else
      call check(hbi_fv_set_break(fv, tout, 1d-2, 0d0), 'hbi_fv_set_break')
    end if
    rc = hbi_fv_step(fv, int(nb, c_int), hist, info)
    if (rc /= 0 .and. rc /= HBI_E_STEP) call check(rc, 'hbi_fv_step')
    call system_clock(c1)
    nrow = int(info%accepted) - kdone
    ! a stop control ends main_fv3dp's loop before the outputs of that step
    stopped = info%status == 1 .or. info%status == 3 .or. info%status == 4
    nout = nrow
    if (stopped) nout = max(nrow - 1, 0)
    do r = 1, nout
      k = kdone + r
      ! output_monitor, main_fv3dp.f90:951-955
      write(52, '(i7,f19.4,7e16.5,f16.4,i6)') k, hist(10 * r - 9), hist(10 * r - 8), hist(10 * r - 7), &
        hist(10 * r - 6), hist(10 * r - 5), hist(10 * r - 4), hist(10 * r - 2), hist(10 * r - 1), &
        real(c1 - c0, 8) / real(crate, 8), int(hist(10 * r))
    end do
    if (nout > 0) call write_local()
    kdone = int(info%accepted)
    if (nout > 0) then
      ! the fields of the last step: output and event steps are always the last of a call
      call check(hbi_fv_get_state(fv, y, pf, disp, vel, qflow, info), 'hbi_fv_get_state')
      call unpermute_fields()
      outfield = interval > 0 .and. mod(kdone, max(interval, 1)) == 0
      if (info%t > tout) then
        outfield = .true.
        tout = info%t + dtout * yr
      end if
      if (outfield .and. nout == nrow) call write_fields()
      vnow = 10d0 ** hist(10 * nout - 8)
      ! event list, main_fv3dp.f90:871-905
      if (.not. slipping .and. vnow > 1d-2 .and. nout == nrow) then
        slipping = .true.
        eventcount = eventcount + 1
        meandisp0 = hist(10 * nout - 7)
        idisp = disp
        hypoloc = maxloc(abs(vel), 1)
        onset_time = info%t
      else if (slipping .and. vnow < 1d-3 .and. nout == nrow) then
        slipping = .false.
        moment = hist(10 * nout - 7) - meandisp0
        write(51, '(i0,i7,f17.2,i7,e15.6,f14.4)') eventcount, kdone, onset_time, hypoloc, moment, &
          (log10(moment * rigid * sg) + 5.9) / 1.5
        write(88) disp - idisp
      end if
    end if
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
  do i = 81, 88
    close(i)
  end do
  close(50); close(51); close(52)
  do i = 1, nl
    close(punit(i))
  end do
  rc = hbi_fv_destroy(fv)
  lrtrn = HACApK_free_leafmtxp(st_leafmtxp_s)
  lrtrn = HACApK_finalize(st_ctl)

contains

  subroutine check(code, what)
    integer(c_int), intent(in) :: code
    character(len=*), intent(in) :: what
    if (code /= 0) then
      write(*, *) 'ERROR: ', what, ': ', trim(hbi_error_text())
      stop 2
    end if
  end subroutine check

  ! output_local rows (main_fv3dp.f90:956-960) of the last call: k, t,
  ! log10 V, disp, sigma_e, tau, pf, mu, psi, kp (the device row holds V,
  ! disp, sigma_e, tau, pf, sigma, mu, psi, kp)
  subroutine write_local()
    integer :: jj, rr, o
    call check(hbi_fv_get_probes(fv, nprow, prow), 'hbi_fv_get_probes')
    do rr = 1, min(int(nprow), nout)
      do jj = 1, nl
        o = 9 * (nl * (rr - 1) + jj - 1)
        write(punit(jj), '(i7,f19.4,8e16.6)') kdone + rr, hist(10 * rr - 9), log10(prow(o + 1)), prow(o + 2), &
          prow(o + 3), prow(o + 4), prow(o + 5), prow(o + 7), prow(o + 8), prow(o + 9)
      end do
    end do
  end subroutine write_local

  ! a mesh-order array in cluster order
  function perm(v) result(p)
    real(8), intent(in) :: v(:)
    real(8) :: p(size(v))
    integer :: jj
    do jj = 1, size(v)
      p(jj) = v(st_sum%lodc(jj))
    end do
  end function perm

  ! the state y (mesh order) in cluster order
  function permute_state() result(yc)
    real(8) :: yc(4 * ncellg)
    integer :: jj, m
    do jj = 1, ncellg
      m = st_sum%lodc(jj)
      yc(4 * jj - 3:4 * jj) = y(4 * m - 3:4 * m)
    end do
  end function permute_state

  ! the fields of hbi_fv_get_state (cluster order) back in mesh order
  subroutine unpermute_fields()
    real(8), allocatable :: t4(:), t1(:)
    integer :: jj, m
    t4 = y
    do jj = 1, ncellg
      m = st_sum%lodc(jj)
      y(4 * m - 3:4 * m) = t4(4 * jj - 3:4 * jj)
    end do
    t1 = pf
    pf(st_sum%lodc) = t1
    t1 = disp
    disp(st_sum%lodc) = t1
    t1 = vel
    vel(st_sum%lodc) = t1
    qflow = q0   ! main_fv3dp keeps qflow = q0 (588)
  end subroutine unpermute_fields

  ! the fields of an output step (main_fv3dp.f90:836-861) and its time row
  subroutine write_fields()
    write(50, '(i7,f19.4)') kdone, info%t
    write(81) vel
    write(82) disp
    write(83) y(3::4)
    write(84) y(2::4)
    write(85) y(4::4) * exp(-y(3::4) / s0)   ! kp after a step, main_fv3dp.f90:787
    write(86) pf
    write(87) qflow
  end subroutine write_fields

  subroutine open_stream(un, name)
    integer, intent(in) :: un
    character(len=*), intent(in) :: name
    character(len=256) :: path
    write(path, '("output/",a,i0,".dat")') name, number
    open(un, file=path, form='unformatted', access='stream', status='replace')
  end subroutine open_stream

end program lhbiem_fv3dp_gpu
