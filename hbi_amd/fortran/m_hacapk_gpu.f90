! m_hacapk_gpu.f90 -- the HACApK call sites of HBI's driver, backed by libhbi.
!
! Same procedure names and argument roles as the reference's HACApK layer so a
! driver written against it (main_LH.f90:590-716, 2111-2122) compiles unchanged
! in its calls; the derived types keep the fields the driver touches
! (st_HACApK_calc_entry geometry pointers, st_ctl%param, st_vec%vs / %lodc),
! and the matrix itself lives on the GPU behind an opaque handle:
!   HACApK_init                m_HACApK_base.f90:246
!   HACApK_generate            m_HACApK_use.f90:133   (GPU build: partition, dense fill, ACA+)
!   HACApK_construct_LH        m_HACApK_use.f90:263   (the device layout is already lattice-packed)
!   HACApK_gen_lattice_vector  m_HACApK_base.f90:3449 (np=1: lodc = cluster permutation)
!   HACApK_adot_lattice_hyp    m_HACApK_solve.f90:220 (y = H x on the GPU)
!   HACApK_free_leafmtxp, HACApK_finalize
! Errors never stop inside the library: the wrappers print hbi_last_error()
! and return a nonzero lrtrn so the driver decides (the reference stops).
module m_HACApK_gpu
  use iso_c_binding
  use m_hbi_abi
  implicit none

  type :: st_HACApK_calc_entry
    character(128) :: problem = ' ', v = 's', md = 's'
    integer :: nd = 0
    real(8) :: w = 0d0
    real(8), pointer :: xcol(:) => null(), ycol(:) => null(), zcol(:) => null()
    real(8), pointer :: xel(:) => null(), xer(:) => null()
    real(8), pointer :: ang(:) => null(), angd(:) => null(), dsl(:) => null(), dsd(:) => null(), rake(:) => null()
    real(8), pointer :: xs1(:) => null(), xs2(:) => null(), xs3(:) => null()
    real(8), pointer :: ys1(:) => null(), ys2(:) => null(), ys3(:) => null()
    real(8), pointer :: zs1(:) => null(), zs2(:) => null(), zs3(:) => null()
    real(8), pointer :: ev11(:) => null(), ev12(:) => null(), ev13(:) => null()
    real(8), pointer :: ev21(:) => null(), ev22(:) => null(), ev23(:) => null()
    real(8), pointer :: ev31(:) => null(), ev32(:) => null(), ev33(:) => null()
    real(8), pointer :: yel(:) => null(), yer(:) => null(), xs4(:) => null(), zs4(:) => null(), ds(:) => null()
  end type st_HACApK_calc_entry

  type :: st_HACApK_lcontrol
    real(8) :: param(100) = 0d0
    integer :: lpmd(50) = 0
    real(8) :: time(10) = 0d0
  end type st_HACApK_lcontrol

  type :: st_HACApK_leafmtxp
    type(c_ptr) :: h = c_null_ptr
    integer :: nd = 0
  end type st_HACApK_leafmtxp

  type :: st_HACApK_LHp
    type(c_ptr) :: h = c_null_ptr
    integer :: prec = 0
  end type st_HACApK_LHp

  type :: st_HACApK_latticevec
    integer :: ndc = 0
    integer, allocatable :: lodc(:)
    real(8), allocatable :: vs(:)
  end type st_HACApK_latticevec

  ! st_ctl%param indices used beyond the reference's: 77 BF16 margin, 78 precision (HBI_PREC_*)
  integer, parameter :: P_BF16 = 77, P_PREC = 78

contains

  integer function HACApK_init(nd, st_ctl, st_bemv, icomm)
    integer, intent(in) :: nd
    type(st_HACApK_lcontrol) :: st_ctl
    type(st_HACApK_calc_entry) :: st_bemv
    integer, optional :: icomm
    ! the defaults of m_HACApK_base.f90:271-302 that shape the partition and ACA+
    st_ctl%param(:) = 0d0
    st_ctl%param(1) = 1
    st_ctl%param(8) = 20
    st_ctl%param(21) = 15
    st_ctl%param(22) = 1.0
    st_ctl%param(51) = 2.0
    st_ctl%param(52) = 1
    st_ctl%param(53) = 100
    st_ctl%param(60) = 2
    st_ctl%param(61) = 3
    st_ctl%param(63) = 1000
    st_ctl%param(64) = 1
    st_ctl%param(72) = 1.0e-6
    st_ctl%param(P_BF16) = 0d0
    st_ctl%param(P_PREC) = HBI_PREC_F64
    st_ctl%lpmd(:) = 0
    st_ctl%lpmd(2) = 1
    st_bemv%nd = nd
    HACApK_init = 0
  end function HACApK_init

  integer function problem_kind(problem)
    character(*), intent(in) :: problem
    select case (trim(problem))
    case ('2dvs')
      problem_kind = HBI_KIND_2DVS
    case ('2dp')
      problem_kind = HBI_KIND_2DP
    case ('3dp')
      problem_kind = HBI_KIND_3DP
    case ('2dph', '2dnh')
      problem_kind = HBI_KIND_2DPH
    case ('2dn')
      problem_kind = HBI_KIND_2DN
    case ('25d')
      problem_kind = HBI_KIND_25D
    case ('2dpfz')
      problem_kind = HBI_KIND_2DPFZ
    case ('2dn_vector')
      problem_kind = HBI_KIND_2DNV
    case ('2dna')
      problem_kind = HBI_KIND_2DNA
    case ('3dph', '3dhr')
      problem_kind = HBI_KIND_OKADA_HS
    case ('3dnr')
      problem_kind = HBI_KIND_OKADA_FS
    case ('3dnt', '3dn')
      problem_kind = HBI_KIND_TRI_FS
    case ('3dht', '3dh')
      problem_kind = HBI_KIND_TRI_HS
    case default
      problem_kind = 0
    end select
  end function problem_kind

  subroutine set_ptr(cp, a)
    type(c_ptr), intent(out) :: cp
    real(8), pointer :: a(:)
    if (associated(a)) then
      cp = c_loc(a(1))
    else
      cp = c_null_ptr
    end if
  end subroutine set_ptr

  integer function HACApK_generate(st_leafmtxp, st_bemv, st_ctl, coord, ztol)
    type(st_HACApK_leafmtxp) :: st_leafmtxp
    type(st_HACApK_calc_entry), target :: st_bemv
    type(st_HACApK_lcontrol) :: st_ctl
    real(8) :: coord(:, :)
    real(8), intent(in) :: ztol
    type(hbi_geom) :: g
    type(hbi_hmat_opts) :: o
    integer(c_int) :: rc
    st_ctl%param(71) = ztol
    g%kind = problem_kind(st_bemv%problem)
    g%n = st_bemv%nd
    g%vcomp = 0
    if (trim(st_bemv%v) == 'n') g%vcomp = 1
    select case (trim(st_bemv%md))
    case ('st')
      g%md = 1
    case ('dp')
      g%md = 2
    case ('open')
      g%md = 3
    case default
      g%md = 0
    end select
    g%w = st_bemv%w
    call set_ptr(g%xcol, st_bemv%xcol); call set_ptr(g%ycol, st_bemv%ycol); call set_ptr(g%zcol, st_bemv%zcol)
    call set_ptr(g%xel, st_bemv%xel); call set_ptr(g%xer, st_bemv%xer)
    call set_ptr(g%ang, st_bemv%ang); call set_ptr(g%angd, st_bemv%angd)
    call set_ptr(g%dsl, st_bemv%dsl); call set_ptr(g%dsd, st_bemv%dsd); call set_ptr(g%rake, st_bemv%rake)
    call set_ptr(g%xs1, st_bemv%xs1); call set_ptr(g%xs2, st_bemv%xs2); call set_ptr(g%xs3, st_bemv%xs3)
    call set_ptr(g%ys1, st_bemv%ys1); call set_ptr(g%ys2, st_bemv%ys2); call set_ptr(g%ys3, st_bemv%ys3)
    call set_ptr(g%zs1, st_bemv%zs1); call set_ptr(g%zs2, st_bemv%zs2); call set_ptr(g%zs3, st_bemv%zs3)
    call set_ptr(g%ev11, st_bemv%ev11); call set_ptr(g%ev12, st_bemv%ev12); call set_ptr(g%ev13, st_bemv%ev13)
    call set_ptr(g%ev21, st_bemv%ev21); call set_ptr(g%ev22, st_bemv%ev22); call set_ptr(g%ev23, st_bemv%ev23)
    call set_ptr(g%ev31, st_bemv%ev31); call set_ptr(g%ev32, st_bemv%ev32); call set_ptr(g%ev33, st_bemv%ev33)
    call set_ptr(g%yel, st_bemv%yel); call set_ptr(g%yer, st_bemv%yer)
    call set_ptr(g%xs4, st_bemv%xs4); call set_ptr(g%zs4, st_bemv%zs4); call set_ptr(g%ds, st_bemv%ds)
    o%leaf = int(st_ctl%param(21))
    o%eta = st_ctl%param(51)
    o%eps_aca = ztol
    o%kmax = int(st_ctl%param(63))
    o%bf16_margin = st_ctl%param(P_BF16)
    rc = hbi_hmat_build(g, o, st_leafmtxp%h)
    st_leafmtxp%nd = st_bemv%nd
    if (rc /= HBI_OK) write(*, *) 'HACApK_generate: ', trim(hbi_error_text())
    HACApK_generate = rc
  end function HACApK_generate

  integer function HACApK_construct_LH(st_LHp, st_leafmtxp, st_bemv, st_ctl, coord, ztol)
    type(st_HACApK_LHp) :: st_LHp
    type(st_HACApK_leafmtxp) :: st_leafmtxp
    type(st_HACApK_calc_entry) :: st_bemv
    type(st_HACApK_lcontrol) :: st_ctl
    real(8) :: coord(:, :)
    real(8), intent(in) :: ztol
    st_LHp%h = st_leafmtxp%h
    st_LHp%prec = int(st_ctl%param(P_PREC))
    HACApK_construct_LH = 0
    if (.not. c_associated(st_LHp%h)) HACApK_construct_LH = -1
  end function HACApK_construct_LH

  integer function HACApK_gen_lattice_vector(st_vec, st_leafmtxp, st_ctl)
    type(st_HACApK_latticevec) :: st_vec
    type(st_HACApK_leafmtxp) :: st_leafmtxp
    type(st_HACApK_lcontrol) :: st_ctl
    integer(c_int), allocatable :: lod(:)
    integer(c_int) :: rc
    st_vec%ndc = st_leafmtxp%nd
    if (allocated(st_vec%lodc)) deallocate(st_vec%lodc)
    if (allocated(st_vec%vs)) deallocate(st_vec%vs)
    allocate(st_vec%lodc(st_vec%ndc), st_vec%vs(st_vec%ndc), lod(st_vec%ndc))
    rc = hbi_hmat_lod(st_leafmtxp%h, lod)
    st_vec%lodc(:) = lod(:) + 1
    st_vec%vs(:) = 0d0
    HACApK_gen_lattice_vector = rc
  end function HACApK_gen_lattice_vector

  subroutine HACApK_adot_lattice_hyp(st_au, st_LHp, st_ctl, wws, st_u)
    type(st_HACApK_latticevec) :: st_au, st_u
    type(st_HACApK_LHp) :: st_LHp
    type(st_HACApK_lcontrol) :: st_ctl
    real(8) :: wws(:)
    integer(c_int) :: rc
    rc = hbi_hmat_matvec(st_LHp%h, st_u%vs, st_au%vs, int(st_LHp%prec, c_int), 0_c_int)
    if (rc /= HBI_OK) write(*, *) 'HACApK_adot_lattice_hyp: ', trim(hbi_error_text())
  end subroutine HACApK_adot_lattice_hyp

  integer function HACApK_free_leafmtxp(st_leafmtxp)
    type(st_HACApK_leafmtxp) :: st_leafmtxp
    HACApK_free_leafmtxp = hbi_hmat_destroy(st_leafmtxp%h)
    st_leafmtxp%h = c_null_ptr
  end function HACApK_free_leafmtxp

  integer function HACApK_finalize(st_ctl)
    type(st_HACApK_lcontrol) :: st_ctl
    HACApK_finalize = 0
  end function HACApK_finalize

end module m_HACApK_gpu
