! lhbiem_gpu.f90 -- Fortran host of the GPU hot path: `lhbiem_gpu <file.in>`.
!
! Reads the reference's input formats (the `key value` .in file of
! main_LH.f90:182-328 with its defaults 133-180, the column-named parameter
! file 530-581, ASCII STL facets 485-510, the 2-D segment and 3dnr/3dhr
! rectangle geometry files 443-471), generates the meshes of the reference
! (2dp/2dvs 1701-1718, 2dph 1720-1737, 2dn/2dnh/25d 1739-1750, 3dp 1752-1779,
! 3dph 1781-1835, STL frames 1940-1990 for 3dnt and 3dht), builds
! the H-matrix through the HACApK call sites (m_HACApK_gpu) and advances the
! quasi-dynamic earthquake cycle with the device RK45 (hbi_rk_step), writing
! the reference's output files: monitor<n>.dat (main_LH.f90:1489-1493
! columns), time<n>.dat, event<n>.dat, the vel/slip/sigma/tau<n>.dat FP64
! streams (cluster order) and EQslip<n>.dat (main_LH.f90:1339-1407,
! 1518-1631, events and snapshots decided on the device), xyz/ind<n>.dat
! (output_coord), local<n>-<id>.dat for each `outloc id` key (at most 10,
! 1-based mesh-order ids; main_LH.f90:322-324, 1099-1122, 1495-1499, from
! probe rows the device records after every step), and the final `time(s)`
! line.  `restart T` continues a run
! from the snapshot of the last interval step in its monitor file (found
! through time<n>.dat, main_LH.f90:836-957).  `sigmaconst F` adds the
! normal-stress matrix and `opening T` the two tensile-dislocation matrices
! with veln/slipn<n>.dat streams (main_LH.f90:683-716).  `forward T` writes H*1
! (main_LH.f90:2586-2664) and stops.  Keys beyond the reference's: leaf
! (param(21)), precision (f64|f32|mixed), bf16margin, outbatch (steps per
! device batch).
program lhbiem_gpu
  use iso_c_binding
  use m_hbi_abi
  use m_HACApK_gpu
  implicit none
  real(8), parameter :: pi = 4 * datan(1.d0), rigid = 32.04d0
  character(len=256) :: input_file, param, pvalue, fname, dum, precision, evlaw, problem, geofile, parameter_file
  character(len=128) :: param2(20)
  integer :: ios, number, ncellg, imax, jmax, nstep, i, k, ncol, leaf, outbatch, nb, nrec_dummy(1)
  real(8) :: ds0, eps_r, eps_h, vref, velmax, velmin, tmax, dtinit, dtmax, crake, dipangle, maxsig, minsig
  real(8) :: a0, b0, dc0, mu0, vpl, velinit, tauinit, sigmainit, sr, bf16margin, t1, t2, sg
  real(8) :: velth, dtout, dtout_co, x0, dtnxt, dummy(9)
  integer :: interval, kstart, kend, kmon, nsnap_c, ksnap, kt_read
  logical :: backslip, sigmaconst, forward, parameterfromfile, limitsigma, restart, found, opening
  type(st_HACApK_lcontrol) :: st_ctl2, st_ctl3, st_ctl4
  type(st_HACApK_leafmtxp) :: st_leafmtxp_n, st_leafmtxp_s2, st_leafmtxp_n2
  type(st_HACApK_LHp) :: st_LHp_n, st_LHp_s2, st_LHp_n2
  type(c_ptr) :: hn_ptr
  integer :: nf
  type(hbi_rk_out_opts) :: oo
  integer(c_int) :: nsnap, nev
  integer(c_long_long), allocatable :: snap_k(:)
  real(8), allocatable :: snap_t(:), fields(:), eqslip(:)
  type(hbi_rk_event), allocatable :: events(:)
  type(st_HACApK_calc_entry) :: st_bemv
  type(st_HACApK_lcontrol) :: st_ctl
  type(st_HACApK_leafmtxp) :: st_leafmtxp_s
  type(st_HACApK_LHp) :: st_LHp_s
  type(st_HACApK_latticevec) :: st_vel, st_sum
  real(8), allocatable :: coord(:, :), wws(:), values(:, :), y(:), slip(:), vel(:), hist(:), vhist(:)
  real(8), allocatable, target :: xcol(:), ycol(:), zcol(:), xel(:), xer(:), ang(:), angd(:), dsl(:), dsd(:), rake(:)
  real(8), allocatable, target :: xs(:, :), ev(:, :), yel(:), yer(:), xq(:, :), dseg(:)
  real(8) :: fwid, rows(7)
  real(8), allocatable :: ds(:), ag(:), bg(:), dcg(:), f0g(:), taug(:), sigmag(:), velg(:), vplg(:), taudotg(:), sigdotg(:)
  real(8), allocatable :: a(:), b(:), dc(:), f0(:), vplv(:), taudot(:), sigdot(:), area(:)
  type(hbi_rk_opts) :: ro
  type(hbi_rk_info) :: info
  type(c_ptr) :: rk
  integer(c_int) :: rc, nrec
  integer :: lrtrn, i_, kdone
  integer :: noutloc, locid(10), nprobe, punit(10)
  integer(c_int) :: probe_idx(10)
  integer(c_long_long) :: nprow
  real(8), allocatable :: prow(:)
  integer(8) :: c0, c1, crate
  real(8) :: mu_avg

  call get_command_argument(1, input_file)
  open(33, file=input_file, status='old', iostat=ios)
  if (ios /= 0) then
    write(*, *) 'ERROR: Failed to open input file'
    stop 1
  end if
  number = 0
  ! defaults (main_LH.f90:133-180); vref, velth are single-precision literals there
  eps_r = 1d-4; eps_h = 1d-4; vref = 1e-6; velmax = 1d7; velmin = 1d-20; tmax = 1d4; dtinit = 1d0; dtmax = 1e10
  maxsig = 300d0; minsig = 1d0; sigmaconst = .false.; forward = .false.; parameterfromfile = .false.
  limitsigma = .true.; backslip = .true.; evlaw = 'aging'; crake = 0d0; dipangle = 90d0; sr = 0d0
  geofile = 'default'; parameter_file = 'default'; ncol = 0; nstep = 1000; leaf = 15
  precision = 'f64'; bf16margin = 1d0; outbatch = 200
  interval = 0; velth = 1e-2; dtout = 365 * 24 * 3600; dtout_co = 1000.0; restart = .false.; opening = .false.
  fwid = 1e8
  noutloc = 0; locid = 0
  a0 = 0.01d0; b0 = 0.015d0; dc0 = 0.01d0; mu0 = 0.6d0; vpl = 1d-9; velinit = 1d-9; tauinit = 0d0; sigmainit = 50d0
  imax = 0; jmax = 0; ncellg = 0; ds0 = 0d0; problem = ' '
  do
    read(33, *, iostat=ios) param, pvalue
    if (ios /= 0) exit
    select case (trim(param))
    case ('problem'); read(pvalue, *) problem
    case ('ncellg'); read(pvalue, *) ncellg
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
    case ('f0'); read(pvalue, *) mu0
    case ('vref'); read(pvalue, *) vref
    case ('vpl'); read(pvalue, *) vpl
    case ('sr'); read(pvalue, *) sr
    case ('velinit'); read(pvalue, *) velinit
    case ('tauinit'); read(pvalue, *) tauinit
    case ('sigmainit'); read(pvalue, *) sigmainit
    case ('dtinit'); read(pvalue, *) dtinit
    case ('dtmax'); read(pvalue, *) dtmax
    case ('tmax'); read(pvalue, *) tmax
    case ('eps_r'); read(pvalue, *) eps_r
    case ('eps_h'); read(pvalue, *) eps_h
    case ('backslip'); read(pvalue, *) backslip
    case ('limitsigma'); read(pvalue, *) limitsigma
    case ('sigmaconst'); read(pvalue, *) sigmaconst
    case ('forward'); read(pvalue, *) forward
    case ('maxsig'); read(pvalue, *) maxsig
    case ('minsig'); read(pvalue, *) minsig
    case ('crake'); read(pvalue, *) crake
    case ('dipangle'); read(pvalue, *) dipangle
    case ('geometry_file'); read(pvalue, '(a)') geofile
    case ('parameterfromfile'); read(pvalue, *) parameterfromfile
    case ('parameter_file'); read(pvalue, '(a)') parameter_file
    case ('parameter_file_ncol'); read(pvalue, *) ncol
    case ('evlaw'); read(pvalue, '(a)') evlaw
    case ('leaf'); read(pvalue, *) leaf
    case ('precision'); read(pvalue, '(a)') precision
    case ('bf16margin'); read(pvalue, *) bf16margin
    case ('outbatch'); read(pvalue, *) outbatch
    case ('interval'); read(pvalue, *) interval
    case ('velth'); read(pvalue, *) velth
    case ('dtout'); read(pvalue, *) dtout
    case ('dtout_co'); read(pvalue, *) dtout_co
    case ('restart'); read(pvalue, *) restart
    case ('opening'); read(pvalue, *) opening
    case ('fwid'); read(pvalue, *) fwid
    case ('outloc')
      if (noutloc == 10) then
        write(*, *) 'ERROR: more than 10 outloc keys'
        stop 1
      end if
      noutloc = noutloc + 1
      read(pvalue, *) locid(noutloc)
    case default
      write(*, *) 'WARNING: ', trim(param), ' is an unknown parameter'
    end select
  end do
  close(33)
  tmax = tmax * 365 * 24 * 3600
  if (interval == 0) interval = nstep
  if (parameter_file == 'default') write(parameter_file, '("examples/param",i0,".dat")') number

  ! ---------------------------------------------------------------- mesh --
  select case (trim(problem))
  case ('3dph', '3dp')
    ncellg = imax * jmax
  case ('3dnt', '3dht')
    ncellg = count_stl_facets(geofile)
  end select
  if (ncellg == 0) then
    write(*, *) 'ERROR: Ncellg is zero'
    stop 1
  end if
  allocate(xcol(ncellg), ycol(ncellg), zcol(ncellg), ds(ncellg), rake(ncellg))
  xcol = 0d0; ycol = 0d0; zcol = 0d0
  select case (trim(problem))
  case ('2dp', '2dvs')
    allocate(xel(ncellg), xer(ncellg))
    do i = 1, ncellg
      ds(i) = ds0
      xel(i) = (i - 1) * ds0
      xer(i) = i * ds0
      xcol(i) = 0.5d0 * (xel(i) + xer(i))
    end do
  case ('2dph', '2dn', '2dnh', '25d')
    ! dipping from the surface (coordinate2dph, main_LH.f90:1720-1737) or
    ! segments from the geometry file (main_LH.f90:443-452, coordinate2dn 1739-1750)
    allocate(xel(ncellg), xer(ncellg), yel(ncellg), yer(ncellg), ang(ncellg), dseg(ncellg))
    if (trim(problem) == '2dph') then
      do i = 1, ncellg
        xel(i) = (i - 1) * ds0 * cos(dipangle * pi / 180)
        xer(i) = i * ds0 * cos(dipangle * pi / 180)
        yel(i) = (i - 1) * ds0 * sin(dipangle * pi / 180)
        yer(i) = i * ds0 * sin(dipangle * pi / 180)
      end do
    else
      open(20, file=geofile, status='old', iostat=ios)
      if (ios /= 0) then
        write(*, *) 'ERROR: Failed to open geometry file'
        stop 1
      end if
      do i = 1, ncellg
        read(20, *) xel(i), xer(i), yel(i), yer(i)
      end do
      close(20)
    end if
    do i = 1, ncellg
      ds(i) = sqrt((xer(i) - xel(i))**2 + (yer(i) - yel(i))**2)
      if (trim(problem) == '2dph') ds(i) = ds0
      ang(i) = datan2(yer(i) - yel(i), xer(i) - xel(i))
      xcol(i) = 0.5d0 * (xel(i) + xer(i))
      ycol(i) = 0.5d0 * (yel(i) + yer(i))
    end do
    dseg = ds
  case ('3dp')
    ! coordinate3dp, main_LH.f90:1752-1779: squares in the x-z plane
    allocate(xq(ncellg, 8))
    do i = 1, imax
      do k = 1, jmax
        i_ = (i - 1) * jmax + k
        xcol(i_) = (i - imax / 2 - 0.5d0) * ds0
        zcol(i_) = -(k - jmax / 2 - 0.5d0) * ds0
        xq(i_, 1) = xcol(i_) + 0.5d0 * ds0; xq(i_, 2) = xcol(i_) - 0.5d0 * ds0
        xq(i_, 3) = xcol(i_) - 0.5d0 * ds0; xq(i_, 4) = xcol(i_) + 0.5d0 * ds0
        xq(i_, 5) = zcol(i_) + 0.5d0 * ds0; xq(i_, 6) = zcol(i_) + 0.5d0 * ds0
        xq(i_, 7) = zcol(i_) - 0.5d0 * ds0; xq(i_, 8) = zcol(i_) - 0.5d0 * ds0
      end do
    end do
    ds = ds0 * ds0
  case ('3dnr', '3dhr')
    ! rectangles from the geometry file, main_LH.f90:460-471
    allocate(ang(ncellg), angd(ncellg), dsl(ncellg), dsd(ncellg))
    open(20, file=geofile, status='old', iostat=ios)
    if (ios /= 0) then
      write(*, *) 'ERROR: Failed to open geometry file'
      stop 1
    end if
    do i = 1, ncellg
      read(20, *) rows
      xcol(i) = rows(1); ycol(i) = rows(2); zcol(i) = rows(3); ang(i) = rows(4); angd(i) = rows(5)
      dsl(i) = rows(6); dsd(i) = rows(7)
      ds(i) = dsl(i) * dsd(i)
    end do
    close(20)
    ds0 = minval(dsl)
  case ('3dph')
    allocate(ang(ncellg), angd(ncellg), dsl(ncellg), dsd(ncellg))
    call mesh_3ddip(imax, jmax, ds0, dipangle, xcol, ycol, zcol, ang, angd)
    ds = ds0 * ds0
    dsl = ds0
    dsd = ds0
  case ('3dnt', '3dht')
    allocate(xs(ncellg, 9), ev(ncellg, 9))
    call read_stl(geofile, ncellg, xs, xcol, ycol, zcol)
    call evcalc(ncellg, xs, ev, ds)
  case default
    write(*, *) 'ERROR: unsupported problem ', trim(problem)
    stop 1
  end select
  rake = crake
  allocate(ag(ncellg), bg(ncellg), dcg(ncellg), f0g(ncellg), taug(ncellg), sigmag(ncellg), velg(ncellg), vplg(ncellg))
  allocate(taudotg(ncellg), sigdotg(ncellg))
  ag = a0; bg = b0; dcg = dc0; f0g = mu0; vplg = vpl; velg = velinit; sigmag = sigmainit; taug = tauinit
  taudotg = 0d0; sigdotg = 0d0
  if (.not. backslip) taudotg = sr
  if (parameterfromfile) then
    open(99, file=parameter_file, status='old', iostat=ios)
    if (ios /= 0) then
      write(*, *) 'ERROR: Failed to open parameter file'
      stop 1
    end if
    read(99, *) param2(1:ncol)
    allocate(values(ncellg, ncol))
    do i = 1, ncellg
      read(99, *) values(i, 1:ncol)
    end do
    close(99)
    do k = 1, ncol
      select case (trim(param2(k)))
      case ('rake'); rake = values(:, k)
      case ('a'); ag = values(:, k)
      case ('b'); bg = values(:, k)
      case ('dc'); dcg = values(:, k)
      case ('f0'); f0g = values(:, k)
      case ('tau'); taug = values(:, k)
      case ('sigma'); sigmag = values(:, k)
      case ('vel'); velg = values(:, k)
      case ('vpl'); vplg = values(:, k)
      case ('taudot'); taudotg = values(:, k)
      case ('sigmadot'); sigdotg = values(:, k)
      end select
    end do
  end if
  rake = rake / 180d0 * pi

  ! ------------------------------------------------------------ H-matrix --
  lrtrn = HACApK_init(ncellg, st_ctl, st_bemv)
  st_ctl%param(21) = leaf
  st_ctl%param(P_BF16) = bf16margin
  select case (trim(precision))
  case ('f32')
    st_ctl%param(P_PREC) = HBI_PREC_F32
  case ('mixed')
    st_ctl%param(P_PREC) = HBI_PREC_MIXED
  case default
    st_ctl%param(P_PREC) = HBI_PREC_F64
  end select
  st_bemv%problem = problem
  st_bemv%xcol => xcol; st_bemv%ycol => ycol; st_bemv%zcol => zcol; st_bemv%rake => rake
  select case (trim(problem))
  case ('2dp', '2dvs')
    st_bemv%xel => xel; st_bemv%xer => xer
  case ('2dph', '2dn', '2dnh', '25d')
    st_bemv%xel => xel; st_bemv%xer => xer; st_bemv%yel => yel; st_bemv%yer => yer
    st_bemv%ang => ang; st_bemv%ds => dseg
    st_bemv%w = fwid
  case ('3dp')
    st_bemv%xs1 => xq(:, 1); st_bemv%xs2 => xq(:, 2); st_bemv%xs3 => xq(:, 3); st_bemv%xs4 => xq(:, 4)
    st_bemv%zs1 => xq(:, 5); st_bemv%zs2 => xq(:, 6); st_bemv%zs3 => xq(:, 7); st_bemv%zs4 => xq(:, 8)
  case ('3dnr', '3dhr')
    st_bemv%ang => ang; st_bemv%angd => angd; st_bemv%dsl => dsl; st_bemv%dsd => dsd
    st_bemv%w = ds0
  case ('3dph')
    st_bemv%ang => ang; st_bemv%angd => angd; st_bemv%dsl => dsl; st_bemv%dsd => dsd
    st_bemv%w = ds0
  case ('3dnt', '3dht')
    st_bemv%xs1 => xs(:, 1); st_bemv%xs2 => xs(:, 2); st_bemv%xs3 => xs(:, 3)
    st_bemv%ys1 => xs(:, 4); st_bemv%ys2 => xs(:, 5); st_bemv%ys3 => xs(:, 6)
    st_bemv%zs1 => xs(:, 7); st_bemv%zs2 => xs(:, 8); st_bemv%zs3 => xs(:, 9)
    st_bemv%ev11 => ev(:, 1); st_bemv%ev12 => ev(:, 2); st_bemv%ev13 => ev(:, 3)
    st_bemv%ev21 => ev(:, 4); st_bemv%ev22 => ev(:, 5); st_bemv%ev23 => ev(:, 6)
    st_bemv%ev31 => ev(:, 7); st_bemv%ev32 => ev(:, 8); st_bemv%ev33 => ev(:, 9)
  end select
  allocate(coord(ncellg, 3))
  coord(:, 1) = xcol; coord(:, 2) = ycol; coord(:, 3) = zcol
  ! main_LH.f90:683-716: the shear matrix; the normal one unless sigma is
  ! constant; with opening the tensile-dislocation matrices too
  select case (trim(problem))
  case ('2dp', '2dn3', '3dp', '2dvs')
    sigmaconst = .true.
  end select
  if (opening) sigmaconst = .false.
  st_bemv%md = 's'
  st_bemv%v = 's'
  call cpu_time(t1)
  lrtrn = HACApK_generate(st_leafmtxp_s, st_bemv, st_ctl, coord, eps_h)
  if (lrtrn /= 0) stop 2
  lrtrn = HACApK_construct_LH(st_LHp_s, st_leafmtxp_s, st_bemv, st_ctl, coord, eps_h)
  if (.not. sigmaconst .and. .not. forward) then
    call build_more(st_ctl2, st_leafmtxp_n, st_LHp_n, 's', 'n')
  end if
  if (opening .and. .not. forward) then
    call build_more(st_ctl3, st_leafmtxp_s2, st_LHp_s2, 'open', 's')
    call build_more(st_ctl4, st_leafmtxp_n2, st_LHp_n2, 'open', 'n')
  end if
  st_bemv%md = 's'
  st_bemv%v = 's'
  allocate(wws(ncellg))
  lrtrn = HACApK_gen_lattice_vector(st_vel, st_leafmtxp_s, st_ctl)
  lrtrn = HACApK_gen_lattice_vector(st_sum, st_leafmtxp_s, st_ctl)
  write(*, *) 'NCELLg', ncellg, ' H-matrix built on the GPU'
  call system('mkdir -p output')

  if (forward) then
    ! H*1 with the okada scaling, main_LH.f90:2641-2664
    st_vel%vs = 1d0
    call HACApK_adot_lattice_hyp(st_sum, st_LHp_s, st_ctl, wws, st_vel)
    select case (trim(problem))
    case ('3dph', '3dnr', '3dhr')
      st_sum%vs = st_sum%vs / ds0
    end select
    write(fname, '("output/stress",i0)') number
    open(29, file=fname)
    do i = 1, ncellg
      i_ = st_sum%lodc(i)
      write(29, '(6e16.4)') xcol(i_), ycol(i_), zcol(i_), 1d0, st_sum%vs(i), 0d0
    end do
    close(29)
    lrtrn = HACApK_free_leafmtxp(st_leafmtxp_s)
    stop
  end if

  ! ----------------------------------------------------- ODE in cluster order --
  allocate(a(ncellg), b(ncellg), dc(ncellg), f0(ncellg), vplv(ncellg), taudot(ncellg), sigdot(ncellg), area(ncellg))
  allocate(y(3 * ncellg), slip(ncellg), vel(ncellg))
  do i = 1, ncellg
    i_ = st_sum%lodc(i)
    a(i) = ag(i_); b(i) = bg(i_); dc(i) = dcg(i_); f0(i) = f0g(i_); vplv(i) = vplg(i_)
    taudot(i) = taudotg(i_); sigdot(i) = sigdotg(i_); area(i) = ds(i_)
    vel(i) = velg(i_)
    y(3 * i - 2) = a(i) * dlog(2 * vref / vel(i) * sinh(taug(i_) / sigmag(i_) / a(i)))
    y(3 * i - 1) = taug(i_)
    y(3 * i) = sigmag(i_)
  end do
  slip = 0d0
  ro%law = 0
  if (trim(evlaw) == 'slip') ro%law = 1
  ro%backslip = merge(1, 0, backslip)
  ro%limitsigma = merge(1, 0, limitsigma)
  ro%prec = int(st_ctl%param(P_PREC))
  ro%vref = vref; ro%minsig = minsig; ro%maxsig = maxsig; ro%dtmax = dtmax
  ro%hdiv = 0d0
  select case (trim(problem))
  case ('3dph', '3dnr', '3dhr')
    ro%hdiv = ds0
  end select
  ro%eps_r = eps_r; ro%velmin = velmin; ro%velmax = velmax
  hn_ptr = c_null_ptr
  if (.not. sigmaconst) hn_ptr = st_LHp_n%h
  rc = hbi_rk_create(st_LHp_s%h, hn_ptr, ro, a, b, dc, f0, vplv, taudot, sigdot, area, rk)
  if (rc /= HBI_OK) then
    write(*, *) 'hbi_rk_create: ', trim(hbi_error_text())
    stop 2
  end if
  if (opening) then
    rc = hbi_rk_set_opening(rk, st_LHp_s2%h, st_LHp_n2%h)
    if (rc /= HBI_OK) then
      write(*, *) 'hbi_rk_set_opening: ', trim(hbi_error_text())
      stop 2
    end if
  end if
  nf = merge(6, 4, opening)
  sg = sum(area)
  x0 = 0d0
  dtnxt = dtinit
  kstart = 1
  kend = 0
  if (restart) then
    ! main_LH.f90:836-957: the last interval step of the monitor file, its time
    ! and next step; the state from that step's snapshot (located through the
    ! time file, whose rows index the snapshot streams)
    write(fname, '("output/monitor",i0,".dat")') number
    open(52, file=fname, status='old')
    do
      read(52, *, iostat=ios) kmon, dummy(1:9)
      if (ios /= 0) exit
      if (mod(kmon, interval) == 0) then
        kstart = kmon + 1
        x0 = dummy(1)
        dtnxt = min(2 * dummy(8), 0.9 * dummy(8) * (dummy(7)**(-0.2)))
      end if
      kend = kmon
    end do
    close(52)
    write(fname, '("output/time",i0,".dat")') number
    open(50, file=fname, status='old')
    found = .false.
    nsnap_c = 0
    do
      read(50, *, iostat=ios) kt_read
      if (ios /= 0) exit
      nsnap_c = nsnap_c + 1
      if (kt_read == kstart - 1) then
        ksnap = nsnap_c
        found = .true.
      end if
    end do
    close(50)
    if (.not. found) then
      write(*, *) 'ERROR: no snapshot of step', kstart - 1
      stop 1
    end if
    call read_snapshot('vel', ksnap, vel)
    call read_snapshot('slip', ksnap, slip)
    call read_snapshot('sigma', ksnap, y(3::3))
    call read_snapshot('tau', ksnap, y(2::3))
    do i = 1, ncellg
      y(3 * i - 2) = a(i) * dlog(2 * vref / vel(i) * sinh(y(3 * i - 1) / y(3 * i) / a(i)))
    end do
    write(*, *) kstart, x0, dtnxt
  end if
  rc = hbi_rk_set_state(rk, y, slip, vel, x0, dtnxt)
  ! device output decisions (main_LH.f90:1188-1195, 1339-1407)
  oo%interval = interval
  oo%capacity = min(outbatch, 64)
  oo%ev_capacity = 16
  oo%k0 = kstart - 1
  oo%velth = velth
  oo%dtout_inter = dtout * 365 * 24 * 3600
  oo%dtout_co = dtout_co
  rc = hbi_rk_set_output(rk, oo)
  allocate(snap_k(oo%capacity), snap_t(oo%capacity), fields(nf * int(oo%capacity) * ncellg))
  allocate(events(oo%ev_capacity), eqslip(int(oo%ev_capacity) * ncellg))
  write(fname, '("output/monitor",i0,".dat")') number
  if (restart) then
    open(52, file=fname, status='old', position='append')
  else
    open(52, file=fname)
  end if
  write(fname, '("output/time",i0,".dat")') number
  call open_text(50, fname, restart)
  write(fname, '("output/event",i0,".dat")') number
  call open_text(44, fname, restart)
  call open_stream(80, 'vel', restart)
  call open_stream(81, 'slip', restart)
  call open_stream(82, 'sigma', restart)
  call open_stream(83, 'tau', restart)
  call open_stream(84, 'EQslip', restart)
  if (opening) then
    call open_stream(85, 'veln', restart)
    call open_stream(86, 'slipn', restart)
  end if
  if (.not. restart) call output_coord()
  ! local output (main_LH.f90:1099-1122): each outloc id (mesh order) found
  ! among the cells is probed at its cluster-order index; the file is opened
  ! afresh, as there, also on restart (rows of steps k > kend only)
  nprobe = 0
  do k = 1, noutloc
    do i = 1, ncellg
      if (st_sum%lodc(i) == locid(k)) then
        nprobe = nprobe + 1
        probe_idx(nprobe) = i - 1
        punit(nprobe) = 53 + k
        write(*, *) xcol(locid(k)), ycol(locid(k)), zcol(locid(k))
        write(fname, '("output/local",i0,"-",i0,".dat")') number, locid(k)
        open(53 + k, file=fname, status='replace')
        exit
      end if
    end do
  end do
  rc = hbi_rk_set_probes(rk, int(nprobe, c_int), probe_idx)
  allocate(prow(8 * max(nprobe, 1) * outbatch))
  allocate(hist(9 * outbatch), vhist(1))
  nrec = 0
  nrec_dummy = 0
  kdone = kstart - 1
  call system_clock(c0, crate)
  if (.not. restart) then
    ! initial row (main_LH.f90:1175-1181): k = 0, errmax = dt = 0
    mu_avg = 0d0
    do i = 1, ncellg
      mu_avg = mu_avg + y(3 * i - 1) / y(3 * i) * area(i)
    end do
    write(52, '(i7,f19.4,7e16.5,i4,f16.4)') 0, 0d0, log10(maxval(abs(vel))), 0d0, mu_avg / sg, &
      maxval(y(3::3)), minval(y(3::3)), 0d0, 0d0, 0, 0d0
  end if
  do while (kdone < nstep)
    nb = min(outbatch, nstep - kdone)
    rc = hbi_rk_step(rk, int(nb, c_int), hist, nrec, nrec_dummy, vhist, info)
    call system_clock(c1)
    ! rows as output_monitor (main_LH.f90:1489-1493); the wall time is the
    ! elapsed time at the end of the device batch the step belongs to.  A
    ! restarted run writes the steps after the old file's last row only.
    do k = 1, int(info%accepted) + kstart - 1 - kdone
      if (kdone + k > kend) &
        write(52, '(i7,f19.4,7e16.5,i4,f16.4)') kdone + k, hist(9 * k - 8), hist(9 * k - 4), hist(9 * k - 3), &
          hist(9 * k - 2), hist(9 * k - 1), hist(9 * k), hist(9 * k - 6), hist(9 * k - 7), int(hist(9 * k - 5)), &
          real(c1 - c0, 8) / real(crate, 8)
    end do
    if (nprobe > 0) call write_local()
    kdone = int(info%accepted) + kstart - 1
    call write_outputs()
    if (rc == HBI_E_STEP) then
      write(*, *) 'ERROR: Runge-Kutta method did not converge'
      exit
    end if
    if (info%status == 1) then
      write(*, *) 'slip rate outside [velmin, velmax] at time step=', kdone
      exit
    end if
    if (info%t > tmax) then
      write(*, *) 'time exceeds tmax at time step=', kdone
      exit
    end if
  end do
  call cpu_time(t2)
  close(52)
  close(50)
  close(44)
  do i = 80, 84
    close(i)
  end do
  if (opening) then
    close(85)
    close(86)
  end if
  do k = 1, nprobe
    close(punit(k))
  end do
  rc = hbi_rk_get_state(rk, y, slip, vel, info)
  write(*, *) 'time(s)', t2 - t1, ' steps', kdone, ' rejected', info%rejected, ' t_end', info%t
  rc = hbi_rk_destroy(rk)
  lrtrn = HACApK_free_leafmtxp(st_leafmtxp_s)
  lrtrn = HACApK_finalize(st_ctl)

contains

  ! rows of output_local (main_LH.f90:1495-1499) for the steps of the last
  ! batch: k, t, log10 V, slip, tau, sigma, mu, psi, vflow (0: no viscous
  ! flow here), veln, slipn
  subroutine write_local()
    integer :: j, r, o
    rc = hbi_rk_get_probes(rk, nprow, prow)
    do r = 1, int(nprow)
      if (kdone + r <= kend) cycle
      do j = 1, nprobe
        o = 8 * (nprobe * (r - 1) + j - 1)
        write(punit(j), '(i7,f19.4,9e16.5)') kdone + r, hist(9 * r - 8), log10(prow(o + 1)), prow(o + 2), &
          prow(o + 3), prow(o + 4), prow(o + 5), prow(o + 6), 0d0, prow(o + 7), prow(o + 8)
      end do
    end do
  end subroutine write_local

  ! the decided outputs of the last batch: time rows and field snapshots
  ! (main_LH.f90:1410-1425, output_field), event rows and coseismic slip
  ! (1376-1379, output_EQslip); a restarted run skips steps it already wrote
  subroutine write_outputs()
    integer :: j, e, n
    n = ncellg
    rc = hbi_rk_get_output(rk, nsnap, snap_k, snap_t, fields, nev, events, eqslip)
    do e = 1, nev
      if (events(e)%k > kend) then
        write(44, '(i0,i8,f17.2,f14.4,i8)') events(e)%count, events(e)%k, events(e)%onset_time, events(e)%mw, &
          events(e)%hypoloc
        write(84) eqslip((e - 1) * n + 1:e * n)
      end if
    end do
    do j = 1, nsnap
      if (snap_k(j) > kend) then
        write(50, '(i7,f19.4)') snap_k(j), snap_t(j)
        write(80) fields((j - 1) * nf * n + 1:(j - 1) * nf * n + n)
        write(81) fields((j - 1) * nf * n + n + 1:(j - 1) * nf * n + 2 * n)
        write(82) fields((j - 1) * nf * n + 2 * n + 1:(j - 1) * nf * n + 3 * n)
        write(83) fields((j - 1) * nf * n + 3 * n + 1:(j - 1) * nf * n + 4 * n)
        if (opening) then
          write(85) fields((j - 1) * nf * n + 4 * n + 1:(j - 1) * nf * n + 5 * n)
          write(86) fields((j - 1) * nf * n + 5 * n + 1:j * nf * n)
        end if
      end if
    end do
  end subroutine write_outputs

  ! another H-matrix of the same mesh: slip mode md onto traction v
  ! (main_LH.f90:698-716)
  subroutine build_more(ctl, leafmtxp, lhp, md, v)
    type(st_HACApK_lcontrol), intent(inout) :: ctl
    type(st_HACApK_leafmtxp), intent(inout) :: leafmtxp
    type(st_HACApK_LHp), intent(inout) :: lhp
    character(len=*), intent(in) :: md, v
    lrtrn = HACApK_init(ncellg, ctl, st_bemv)
    ctl%param = st_ctl%param
    st_bemv%md = md
    st_bemv%v = v
    lrtrn = HACApK_generate(leafmtxp, st_bemv, ctl, coord, eps_h)
    if (lrtrn /= 0) stop 2
    lrtrn = HACApK_construct_LH(lhp, leafmtxp, st_bemv, ctl, coord, eps_h)
  end subroutine build_more

  subroutine open_text(u, path, append)
    integer, intent(in) :: u
    character(len=*), intent(in) :: path
    logical, intent(in) :: append
    if (append) then
      open(u, file=path, position='append')
    else
      open(u, file=path, status='replace')
    end if
  end subroutine open_text

  subroutine open_stream(u, name, append)
    integer, intent(in) :: u
    character(len=*), intent(in) :: name
    logical, intent(in) :: append
    character(len=256) :: path
    write(path, '("output/",a,i0,".dat")') name, number
    if (append) then
      open(u, file=path, form='unformatted', access='stream', status='old', position='append')
    else
      open(u, file=path, form='unformatted', access='stream', status='replace')
    end if
  end subroutine open_stream

  ! snapshot j (1-based) of output/<name><n>.dat into v (cluster order)
  subroutine read_snapshot(name, j, v)
    character(len=*), intent(in) :: name
    integer, intent(in) :: j
    real(8), intent(out) :: v(:)
    character(len=256) :: path
    integer :: u
    integer(8) :: pos
    write(path, '("output/",a,i0,".dat")') name, number
    open(newunit=u, file=path, form='unformatted', access='stream', status='old')
    pos = int(j - 1, 8) * int(ncellg, 8) * 8 + 1
    read(u, pos=pos) v
    close(u)
  end subroutine read_snapshot

  ! output_coord (main_LH.f90:1659-1697): collocation points and element
  ! ids in cluster order
  subroutine output_coord()
    integer :: u1, u2, j, jj
    character(len=256) :: path
    write(path, '("output/xyz",i0,".dat")') number
    open(newunit=u1, file=path, status='replace')
    write(path, '("output/ind",i0,".dat")') number
    open(newunit=u2, file=path, status='replace')
    do j = 1, ncellg
      jj = st_sum%lodc(j)
      write(u2, *) jj
      write(u1, '(3e15.6)') xcol(jj), ycol(jj), zcol(jj)
    end do
    close(u1)
    close(u2)
  end subroutine output_coord

  subroutine mesh_3ddip(imax, jmax, ds0, dipangle, xcol, ycol, zcol, ang, angd)
    integer, intent(in) :: imax, jmax
    real(8), intent(in) :: ds0, dipangle
    real(8), intent(out) :: xcol(:), ycol(:), zcol(:), ang(:), angd(:)
    real(8) :: yr(0:jmax), zr(0:jmax)
    integer :: i, j, k
    yr(0) = 0d0; zr(0) = 0d0
    do j = 1, jmax
      yr(j) = yr(j - 1) - ds0 * cos(dipangle * pi / 180d0)
      zr(j) = zr(j - 1) - ds0 * sin(dipangle * pi / 180d0)
    end do
    k = 0
    do i = 1, imax
      do j = 1, jmax
        k = k + 1
        xcol(k) = (i - imax / 2 - 0.5d0) * ds0
        ycol(k) = 0.5d0 * (yr(j - 1) + yr(j))
        zcol(k) = 0.5d0 * (zr(j - 1) + zr(j))
        angd(k) = datan2(zr(j - 1) - zr(j), yr(j - 1) - yr(j))
        ang(k) = 0d0
      end do
    end do
  end subroutine mesh_3ddip

  integer function count_stl_facets(path)
    character(*), intent(in) :: path
    integer :: u, ios, nl
    open(newunit=u, file=path, status='old', iostat=ios)
    if (ios /= 0) then
      write(*, *) 'ERROR: Failed to open geometry file'
      stop 1
    end if
    nl = 0
    do
      read(u, '(a)', iostat=ios)
      if (ios /= 0) exit
      nl = nl + 1
    end do
    close(u)
    count_stl_facets = nl / 7
  end function count_stl_facets

  subroutine read_stl(path, n, xs, xcol, ycol, zcol)
    character(*), intent(in) :: path
    integer, intent(in) :: n
    real(8), intent(out) :: xs(n, 9), xcol(n), ycol(n), zcol(n)
    character(len=64) :: w
    integer :: u, k
    open(newunit=u, file=path, status='old')
    do
      read(u, *) w
      if (w == 'facet') exit
    end do
    do k = 1, n
      read(u, *)
      read(u, *) w, xs(k, 1), xs(k, 4), xs(k, 7)
      read(u, *) w, xs(k, 2), xs(k, 5), xs(k, 8)
      read(u, *) w, xs(k, 3), xs(k, 6), xs(k, 9)
      read(u, *)
      read(u, *)
      if (k < n) read(u, *)
      xcol(k) = (xs(k, 1) + xs(k, 2) + xs(k, 3)) / 3
      ycol(k) = (xs(k, 4) + xs(k, 5) + xs(k, 6)) / 3
      zcol(k) = (xs(k, 7) + xs(k, 8) + xs(k, 9)) / 3
    end do
    close(u)
  end subroutine read_stl

  subroutine evcalc(n, xs, ev, ds)
    integer, intent(in) :: n
    real(8), intent(in) :: xs(n, 9)
    real(8), intent(out) :: ev(n, 9), ds(n)
    real(8) :: vba(3), vca(3), rr, t1, t2, t3
    integer :: k
    do k = 1, n
      vba = (/ xs(k, 2) - xs(k, 1), xs(k, 5) - xs(k, 4), xs(k, 8) - xs(k, 7) /)
      vca = (/ xs(k, 3) - xs(k, 1), xs(k, 6) - xs(k, 4), xs(k, 9) - xs(k, 7) /)
      ev(k, 7) = vba(2) * vca(3) - vba(3) * vca(2)
      ev(k, 8) = vba(3) * vca(1) - vba(1) * vca(3)
      ev(k, 9) = vba(1) * vca(2) - vba(2) * vca(1)
      rr = sqrt(ev(k, 7) * ev(k, 7) + ev(k, 8) * ev(k, 8) + ev(k, 9) * ev(k, 9))
      ev(k, 7:9) = ev(k, 7:9) / rr
      if (abs(ev(k, 9)) < 1.0d0) then
        ev(k, 1) = -ev(k, 8); ev(k, 2) = ev(k, 7); ev(k, 3) = 0d0
        rr = sqrt(ev(k, 1) * ev(k, 1) + ev(k, 2) * ev(k, 2))
        ev(k, 1) = ev(k, 1) / rr; ev(k, 2) = ev(k, 2) / rr
      else
        ev(k, 1) = 1d0; ev(k, 2) = 0d0; ev(k, 3) = 0d0
      end if
      ev(k, 4) = ev(k, 8) * ev(k, 3) - ev(k, 9) * ev(k, 2)
      ev(k, 5) = ev(k, 9) * ev(k, 1) - ev(k, 7) * ev(k, 3)
      ev(k, 6) = ev(k, 7) * ev(k, 2) - ev(k, 8) * ev(k, 1)
      t1 = vba(1) * vba(1) + vba(2) * vba(2) + vba(3) * vba(3)
      t2 = vca(1) * vca(1) + vca(2) * vca(2) + vca(3) * vca(3)
      t3 = vba(1) * vca(1) + vba(2) * vca(2) + vba(3) * vca(3)
      ds(k) = 0.5d0 * sqrt(t1 * t2 - t3 * t3)
    end do
  end subroutine evcalc

end program lhbiem_gpu
