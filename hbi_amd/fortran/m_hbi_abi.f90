! m_hbi_abi.f90 -- ISO_C_BINDING interfaces to libhbi (include/hbi.h).
module m_hbi_abi
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: HBI_OK = 0, HBI_E_STEP = -4, HBI_RK_OUTPUT_FULL = 2
  integer(c_int), parameter :: HBI_KIND_2DVS = 1, HBI_KIND_OKADA_HS = 2, HBI_KIND_OKADA_FS = 3, HBI_KIND_TRI_FS = 4, &
    & HBI_KIND_TRI_HS = 5, HBI_KIND_2DP = 6, HBI_KIND_3DP = 7, HBI_KIND_2DPH = 8, HBI_KIND_2DN = 9, &
    & HBI_KIND_25D = 10, HBI_KIND_2DPFZ = 11, HBI_KIND_2DNV = 12, HBI_KIND_2DNA = 13
  integer(c_int), parameter :: HBI_PREC_F64 = 0, HBI_PREC_F32 = 1, HBI_PREC_MIXED = 2
  integer(c_int), parameter :: HBI_FV_INJECTIONQ = 0, HBI_FV_INJECTIONP = 1, HBI_FV_STEADYSTATE = 2, &
    & HBI_FV_THRUST = 3, HBI_FV_OTHER = 4

  type, bind(C) :: hbi_geom
    integer(c_int) :: kind = 0, n = 0, vcomp = 0, md = 0
    real(c_double) :: w = 0d0
    type(c_ptr) :: xcol = c_null_ptr, ycol = c_null_ptr, zcol = c_null_ptr
    type(c_ptr) :: xel = c_null_ptr, xer = c_null_ptr
    type(c_ptr) :: ang = c_null_ptr, angd = c_null_ptr, dsl = c_null_ptr, dsd = c_null_ptr, rake = c_null_ptr
    type(c_ptr) :: xs1 = c_null_ptr, xs2 = c_null_ptr, xs3 = c_null_ptr
    type(c_ptr) :: ys1 = c_null_ptr, ys2 = c_null_ptr, ys3 = c_null_ptr
    type(c_ptr) :: zs1 = c_null_ptr, zs2 = c_null_ptr, zs3 = c_null_ptr
    type(c_ptr) :: ev11 = c_null_ptr, ev12 = c_null_ptr, ev13 = c_null_ptr
    type(c_ptr) :: ev21 = c_null_ptr, ev22 = c_null_ptr, ev23 = c_null_ptr
    type(c_ptr) :: ev31 = c_null_ptr, ev32 = c_null_ptr, ev33 = c_null_ptr
    type(c_ptr) :: yel = c_null_ptr, yer = c_null_ptr, xs4 = c_null_ptr, zs4 = c_null_ptr, ds = c_null_ptr
  end type hbi_geom

  type, bind(C) :: hbi_hmat_opts
    integer(c_int) :: leaf = 15
    real(c_double) :: eta = 2d0
    real(c_double) :: eps_aca = 1d-4
    integer(c_int) :: kmax = 1000
    real(c_double) :: bf16_margin = 0d0
    integer(c_int) :: tile_rows = 0
    integer(c_int) :: chunk_kb = 0
    integer(c_int) :: dense_launch = 0
  end type hbi_hmat_opts

  type, bind(C) :: hbi_rk_opts
    integer(c_int) :: law = 0, backslip = 1, limitsigma = 1, prec = 0
    real(c_double) :: vref = 1d-6, minsig = 1d0, maxsig = 300d0, dtmax = 1d10
    real(c_double) :: hdiv = 0d0, eps_r = 1d-4, velmin = 1d-20, velmax = 1d7
  end type hbi_rk_opts

  ! fluid-driven cycle of main_fv.f90 (hbi_fv_*); defaults of main_fv.f90:130-162
  type, bind(C) :: hbi_fv_opts
    integer(c_int) :: law = 0, backslip = 0, prec = 0, pressure_dependent = 0, pfconst = 0
    integer(c_int) :: bcl = 0, bcr = 0, setting = 0, outpertime = 0
    real(c_double) :: vref = 1d-6, eps_r = 1d-4, hdiv = 0d0, ds0 = 0d0, kL = 1d8, kpmin = 0d0, s0 = 0d0
    real(c_double) :: eta = 0d0, beta = 0d0, Bs = 0d0, pbcl = 0d0, pbcr = 0d0, q0 = 0d0, pinj = 0d0
    real(c_double) :: tinj_s = 0d0, dtout_s = 0d0, velmin = 1d-16, velmax = 1d7, tmax_s = 1d12
    integer(c_int) :: imax = 0, jmax = 0, bct = 0, bcb = 0
    real(c_double) :: dtmax = 1d10, minsig = 0d0, phi0 = 0d0
  end type hbi_fv_opts

  type, bind(C) :: hbi_fv_info
    real(c_double) :: t, dt_next, dt_last, errmax_last
    integer(c_long_long) :: attempts, accepted, rejected
    real(c_double) :: vmax
    integer(c_int) :: niter, status
  end type hbi_fv_info

  type, bind(C) :: hbi_rk_info
    real(c_double) :: t, dt_next, dt_last, errmax_last
    integer(c_long_long) :: attempts, accepted, rejected
    real(c_double) :: vmax
    integer(c_int) :: status
  end type hbi_rk_info

  type, bind(C) :: hbi_rk_out_opts
    integer(c_int) :: interval = 0, capacity = 64, ev_capacity = 16, pad = 0
    integer(c_long_long) :: k0 = 0
    real(c_double) :: velth = 1d-2, dtout_inter = 3.1536d7, dtout_co = 1d3
  end type hbi_rk_out_opts

  type, bind(C) :: hbi_rk_event
    integer(c_long_long) :: count, k
    real(c_double) :: onset_time, mw, moment
    integer(c_int) :: hypoloc, pad
  end type hbi_rk_event

  interface
    integer(c_int) function hbi_version() bind(C, name="hbi_version")
      import :: c_int
    end function
    type(c_ptr) function hbi_last_error() bind(C, name="hbi_last_error")
      import :: c_ptr
    end function
    integer(c_int) function hbi_hmat_build(g, opts, h) bind(C, name="hbi_hmat_build")
      import :: c_int, c_ptr, hbi_geom, hbi_hmat_opts
      type(hbi_geom), intent(in) :: g
      type(hbi_hmat_opts), intent(in) :: opts
      type(c_ptr), intent(out) :: h
    end function
    integer(c_int) function hbi_hmat_destroy(h) bind(C, name="hbi_hmat_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function hbi_hmat_lod(h, lod) bind(C, name="hbi_hmat_lod")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), intent(out) :: lod(*)
    end function
    integer(c_int) function hbi_hmat_matvec(h, x, y, prec, on_device) bind(C, name="hbi_hmat_matvec")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(out) :: y(*)
      integer(c_int), value :: prec, on_device
    end function
    ! y = H x on device pointers, no allocation (scratch kept by the library on each GPU)
    integer(c_int) function hbi_hmat_matvec_noalloc(h, x, y, prec) bind(C, name="hbi_hmat_matvec_noalloc")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, x, y
      integer(c_int), value :: prec
    end function
    integer(c_int) function hbi_rk_create(hs, hn, opts, a, b, dc, f0, vpl, taudot, sigdot, area, r) &
        bind(C, name="hbi_rk_create")
      import :: c_int, c_ptr, c_double, hbi_rk_opts
      type(c_ptr), value :: hs, hn
      type(hbi_rk_opts), intent(in) :: opts
      real(c_double), intent(in) :: a(*), b(*), dc(*), f0(*), vpl(*), taudot(*), sigdot(*), area(*)
      type(c_ptr), intent(out) :: r
    end function
    integer(c_int) function hbi_rk_destroy(r) bind(C, name="hbi_rk_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: r
    end function
    integer(c_int) function hbi_rk_set_state(r, y, slip, vel, t, dt) bind(C, name="hbi_rk_set_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: r
      real(c_double), intent(in) :: y(*), slip(*), vel(*)
      real(c_double), value :: t, dt
    end function
    integer(c_int) function hbi_rk_get_state(r, y, slip, vel, info) bind(C, name="hbi_rk_get_state")
      import :: c_int, c_ptr, c_double, hbi_rk_info
      type(c_ptr), value :: r
      real(c_double), intent(out) :: y(*), slip(*), vel(*)
      type(hbi_rk_info), intent(out) :: info
    end function
    integer(c_int) function hbi_rk_set_opening(r, hs2, hn2) bind(C, name="hbi_rk_set_opening")
      import :: c_int, c_ptr
      type(c_ptr), value :: r, hs2, hn2
    end function
    integer(c_int) function hbi_rk_set_probes(r, nprobe, idx) bind(C, name="hbi_rk_set_probes")
      import :: c_int, c_ptr
      type(c_ptr), value :: r
      integer(c_int), value :: nprobe
      integer(c_int), intent(in) :: idx(*)
    end function
    integer(c_int) function hbi_rk_get_probes(r, nrows, rows) bind(C, name="hbi_rk_get_probes")
      import :: c_int, c_ptr, c_double, c_long_long
      type(c_ptr), value :: r
      integer(c_long_long), intent(out) :: nrows
      real(c_double), intent(out) :: rows(*)
    end function
    integer(c_int) function hbi_rk_set_output(r, o) bind(C, name="hbi_rk_set_output")
      import :: c_int, c_ptr, hbi_rk_out_opts
      type(c_ptr), value :: r
      type(hbi_rk_out_opts), intent(in) :: o
    end function
    integer(c_int) function hbi_rk_get_output(r, nsnap, snap_k, snap_t, fields, nev, events, eqslip) &
        bind(C, name="hbi_rk_get_output")
      import :: c_int, c_ptr, c_double, c_long_long, hbi_rk_event
      type(c_ptr), value :: r
      integer(c_int), intent(out) :: nsnap, nev
      integer(c_long_long), intent(out) :: snap_k(*)
      real(c_double), intent(out) :: snap_t(*), fields(*), eqslip(*)
      type(hbi_rk_event), intent(out) :: events(*)
    end function
    integer(c_int) function hbi_rk_step(r, nsteps, hist, nrec, rec, vhist, info) bind(C, name="hbi_rk_step")
      import :: c_int, c_ptr, c_double, hbi_rk_info
      type(c_ptr), value :: r
      integer(c_int), value :: nsteps, nrec
      real(c_double), intent(out) :: hist(*)
      integer(c_int), intent(in) :: rec(*)
      real(c_double), intent(out) :: vhist(*)
      type(hbi_rk_info), intent(out) :: info
    end function
    integer(c_int) function hbi_hmat_matmul(h, x, y, nrhs, prec) bind(C, name="hbi_hmat_matmul")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(out) :: y(*)
      integer(c_int), value :: nrhs, prec
    end function
    integer(c_int) function hbi_rk_step_many(rs, m, nsteps, infos) bind(C, name="hbi_rk_step_many")
      import :: c_int, c_ptr, hbi_rk_info
      type(c_ptr), intent(in) :: rs(*)
      integer(c_int), value :: m, nsteps
      type(hbi_rk_info), intent(out) :: infos(*)
    end function
    ! ensembles
    integer(c_int) function hbi_ens_create(hs, opts, nreal, a, b, dc, f0, vpl, taudot, sigdot, e) &
        bind(C, name="hbi_ens_create")
      import :: c_int, c_ptr, c_double, hbi_rk_opts
      type(c_ptr), value :: hs
      type(hbi_rk_opts), intent(in) :: opts
      integer(c_int), value :: nreal
      real(c_double), intent(in) :: a(*), b(*), dc(*), f0(*), vpl(*), taudot(*), sigdot(*)
      type(c_ptr), intent(out) :: e
    end function
    integer(c_int) function hbi_ens_destroy(e) bind(C, name="hbi_ens_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: e
    end function
    integer(c_int) function hbi_ens_set_state(e, y, t, dt) bind(C, name="hbi_ens_set_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: e
      real(c_double), intent(in) :: y(*)
      real(c_double), value :: t, dt
    end function
    integer(c_int) function hbi_ens_set_clock(e, t, dt_next) bind(C, name="hbi_ens_set_clock")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: e
      real(c_double), intent(in) :: t(*), dt_next(*)
    end function
    integer(c_int) function hbi_ens_get_clock(e, t, dt_next) bind(C, name="hbi_ens_get_clock")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: e
      real(c_double), intent(out) :: t(*), dt_next(*)
    end function
    integer(c_int) function hbi_ens_get_state(e, y, slip, vel, t, accepted, rejected, status) &
        bind(C, name="hbi_ens_get_state")
      import :: c_int, c_ptr, c_double, c_long_long
      type(c_ptr), value :: e
      real(c_double), intent(out) :: y(*), slip(*), vel(*), t(*)
      integer(c_long_long), intent(out) :: accepted(*), rejected(*)
      integer(c_int), intent(out) :: status(*)
    end function
    integer(c_int) function hbi_ens_step(e, nsteps, hist, attempts) bind(C, name="hbi_ens_step")
      import :: c_int, c_ptr, c_double, c_long_long
      type(c_ptr), value :: e
      integer(c_int), value :: nsteps
      real(c_double), intent(out) :: hist(*)
      integer(c_long_long), intent(out) :: attempts
    end function
    integer(c_int) function hbi_ens_set_probes(e, nprobe, idx) bind(C, name="hbi_ens_set_probes")
      import :: c_int, c_ptr
      type(c_ptr), value :: e
      integer(c_int), value :: nprobe
      integer(c_int), intent(in) :: idx(*)
    end function
    integer(c_int) function hbi_ens_set_step_grid(e, cap, h) bind(C, name="hbi_ens_set_step_grid")
      import :: c_int, c_ptr, c_double, c_long_long
      type(c_ptr), value :: e
      integer(c_long_long), value :: cap
      real(c_double), intent(in) :: h(*)
    end function
    integer(c_int) function hbi_ens_set_controller(e, mode, m, safety, ratio_lo, ratio_hi) &
        bind(C, name="hbi_ens_set_controller")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: e, m
      integer(c_int), value :: mode
      real(c_double), value :: safety, ratio_lo, ratio_hi
    end function
    ! learned step size
    integer(c_int) function hbi_mlp_create(seed, m) bind(C, name="hbi_mlp_create")
      import :: c_int, c_ptr, c_long_long
      integer(c_long_long), value :: seed
      type(c_ptr), intent(out) :: m
    end function
    integer(c_int) function hbi_mlp_destroy(m) bind(C, name="hbi_mlp_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
    end function
    integer(c_int) function hbi_mlp_set_params(m, params, mu, sd) bind(C, name="hbi_mlp_set_params")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: m
      real(c_float), intent(in) :: params(*), mu(*), sd(*)
    end function
    integer(c_int) function hbi_rk_set_controller(r, mode, m, safety, ratio_lo, ratio_hi) &
        bind(C, name="hbi_rk_set_controller")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: r, m
      integer(c_int), value :: mode
      real(c_double), value :: safety, ratio_lo, ratio_hi
    end function
    integer(c_int) function hbi_fv_create(hs, hn, o, a, b, dc, f0, vpl, kpmaxv, kTv, phi, pfhyd, qin, area, f) &
        bind(C, name="hbi_fv_create")
      import :: c_int, c_ptr, c_double, hbi_fv_opts
      type(c_ptr), value :: hs, hn
      type(hbi_fv_opts), intent(in) :: o
      real(c_double), intent(in) :: a(*), b(*), dc(*), f0(*), vpl(*), kpmaxv(*), kTv(*), phi(*), pfhyd(*)
      real(c_double), intent(in) :: qin(*), area(*)
      type(c_ptr), intent(out) :: f
    end function
    integer(c_int) function hbi_fv_destroy(f) bind(C, name="hbi_fv_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: f
    end function
    integer(c_int) function hbi_fv_set_state(f, y, pf, disp, t, dt) bind(C, name="hbi_fv_set_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: f
      real(c_double), intent(in) :: y(*), pf(*), disp(*)
      real(c_double), value :: t, dt
    end function
    integer(c_int) function hbi_fv_get_state(f, y, pf, disp, vel, qflow, info) bind(C, name="hbi_fv_get_state")
      import :: c_int, c_ptr, c_double, hbi_fv_info
      type(c_ptr), value :: f
      real(c_double), intent(out) :: y(*), pf(*), disp(*), vel(*), qflow(*)
      type(hbi_fv_info), intent(out) :: info
    end function
    integer(c_int) function hbi_fv_set_wells(f, nwell, iw, jw, kleng, qvals, qtimes) bind(C, name="hbi_fv_set_wells")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: f
      integer(c_int), value :: nwell
      integer(c_int), intent(in) :: iw(*), jw(*), kleng(*)
      real(c_double), intent(in) :: qvals(*), qtimes(*)
    end function
    integer(c_int) function hbi_fv_set_probes(f, nprobe, idx) bind(C, name="hbi_fv_set_probes")
      import :: c_int, c_ptr
      type(c_ptr), value :: f
      integer(c_int), value :: nprobe
      integer(c_int), intent(in) :: idx(*)
    end function
    integer(c_int) function hbi_fv_get_probes(f, nrows, rows) bind(C, name="hbi_fv_get_probes")
      import :: c_int, c_ptr, c_double, c_long_long
      type(c_ptr), value :: f
      integer(c_long_long), intent(out) :: nrows
      real(c_double), intent(out) :: rows(*)
    end function
    integer(c_int) function hbi_fv_set_break(f, t, vmax_above, vmax_below) bind(C, name="hbi_fv_set_break")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: f
      real(c_double), value :: t, vmax_above, vmax_below
    end function
    integer(c_int) function hbi_fv_step(f, nsteps, hist, info) bind(C, name="hbi_fv_step")
      import :: c_int, c_ptr, c_double, hbi_fv_info
      type(c_ptr), value :: f
      integer(c_int), value :: nsteps
      real(c_double), intent(out) :: hist(*)
      type(hbi_fv_info), intent(out) :: info
    end function
  end interface

contains

  ! text of hbi_last_error() as a Fortran string
  function hbi_error_text() result(s)
    character(len=512) :: s
    character(kind=c_char), pointer :: p(:)
    integer :: i
    s = ' '
    call c_f_pointer(hbi_last_error(), p, [512])
    do i = 1, 512
      if (p(i) == c_null_char) exit
      s(i:i) = p(i)
    end do
  end function hbi_error_text

end module m_hbi_abi
