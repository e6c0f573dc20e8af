! ref_harness.f90 -- TEST INFRASTRUCTURE ONLY.
!
! A C-callable wrapper around the reference's own matrix-entry function
! HACApK_entry_ij (m_HACApK_calc_entry_ij.f90:24-90), linked against the
! reference kernel sources (m_const.f90, okada.f90, TDstressFS.f90,
! m_HACApK_calc_entry_ij.f90) compiled unmodified by oracle/Makefile into
! oracle/_ref/libhbiref.so.  It exists to pin the oracle's kernel restatement
! and to write the golden vectors under tests/golden/.
!
! geo(n,33) columns: xcol ycol zcol xel xer ang angd dsl dsd rake
!                    xs1 xs2 xs3 ys1 ys2 ys3 zs1 zs2 zs3
!                    ev11 ev12 ev13 ev21 ev22 ev23 ev31 ev32 ev33
!                    yel yer xs4 zs4 ds
! kind: 1 '2dvs', 2 '3dph', 3 '3dnr', 4 '3dnt', 5 '3dht', 6 '2dp', 7 '3dp',
!       8 '2dph', 9 '2dn', 10 '25d', 11 '2dpfz', 12 '2dn_vector', 13 '2dna';
!       vcomp: 0 's', 1 'n';
! md: 0 's', 1 'st', 2 'dp', 3 'open'.  ii, jj are 0-based.
module hbiref_harness
  use iso_c_binding
  use m_HACApK_calc_entry_ij
  use dtriangular
  implicit none
contains
  subroutine hbiref_entries(kind, n, vcomp, md, w, geo, np, ii, jj, out) bind(C, name="hbiref_entries")
    integer(c_int), value :: kind, n, vcomp, md, np
    real(c_double), value :: w
    real(c_double), intent(in) :: geo(n, 33)
    integer(c_int), intent(in) :: ii(np), jj(np)
    real(c_double), intent(out) :: out(np)
    type(st_HACApK_calc_entry) :: bem
    integer :: p

    select case (kind)
    case (1)
      bem%problem = '2dvs'
    case (2)
      bem%problem = '3dph'
    case (3)
      bem%problem = '3dnr'
    case (4)
      bem%problem = '3dnt'
    case (5)
      bem%problem = '3dht'
    case (6)
      bem%problem = '2dp'
    case (7)
      bem%problem = '3dp'
    case (8)
      bem%problem = '2dph'
    case (9)
      bem%problem = '2dn'
    case (10)
      bem%problem = '25d'
    case (11)
      bem%problem = '2dpfz'
    case (12)
      bem%problem = '2dn_vector'
    case (13)
      bem%problem = '2dna'
    end select
    if (vcomp == 1) then
      bem%v = 'n'
    else
      bem%v = 's'
    end if
    select case (md)
    case (1)
      bem%md = 'st'
    case (2)
      bem%md = 'dp'
    case (3)
      bem%md = 'open'
    case default
      bem%md = 's'
    end select
    bem%w = w
    allocate(bem%xcol(n), bem%ycol(n), bem%zcol(n), bem%xel(n), bem%xer(n))
    allocate(bem%ang(n), bem%angd(n), bem%dsl(n), bem%dsd(n), bem%rake(n))
    allocate(bem%xs1(n), bem%xs2(n), bem%xs3(n), bem%ys1(n), bem%ys2(n), bem%ys3(n))
    allocate(bem%zs1(n), bem%zs2(n), bem%zs3(n))
    allocate(bem%ev11(n), bem%ev12(n), bem%ev13(n), bem%ev21(n), bem%ev22(n), bem%ev23(n))
    allocate(bem%ev31(n), bem%ev32(n), bem%ev33(n))
    allocate(bem%yel(n), bem%yer(n), bem%xs4(n), bem%zs4(n), bem%ds(n))
    bem%xcol = geo(:, 1); bem%ycol = geo(:, 2); bem%zcol = geo(:, 3)
    bem%xel = geo(:, 4); bem%xer = geo(:, 5)
    bem%ang = geo(:, 6); bem%angd = geo(:, 7); bem%dsl = geo(:, 8); bem%dsd = geo(:, 9)
    bem%rake = geo(:, 10)
    bem%xs1 = geo(:, 11); bem%xs2 = geo(:, 12); bem%xs3 = geo(:, 13)
    bem%ys1 = geo(:, 14); bem%ys2 = geo(:, 15); bem%ys3 = geo(:, 16)
    bem%zs1 = geo(:, 17); bem%zs2 = geo(:, 18); bem%zs3 = geo(:, 19)
    bem%ev11 = geo(:, 20); bem%ev12 = geo(:, 21); bem%ev13 = geo(:, 22)
    bem%ev21 = geo(:, 23); bem%ev22 = geo(:, 24); bem%ev23 = geo(:, 25)
    bem%ev31 = geo(:, 26); bem%ev32 = geo(:, 27); bem%ev33 = geo(:, 28)
    bem%yel = geo(:, 29); bem%yer = geo(:, 30); bem%xs4 = geo(:, 31); bem%zs4 = geo(:, 32); bem%ds = geo(:, 33)
    do p = 1, np
      out(p) = HACApK_entry_ij(ii(p) + 1, jj(p) + 1, bem)
    end do
    deallocate(bem%xcol, bem%ycol, bem%zcol, bem%xel, bem%xer, bem%ang, bem%angd, bem%dsl, bem%dsd, bem%rake)
    deallocate(bem%xs1, bem%xs2, bem%xs3, bem%ys1, bem%ys2, bem%ys3, bem%zs1, bem%zs2, bem%zs3)
    deallocate(bem%ev11, bem%ev12, bem%ev13, bem%ev21, bem%ev22, bem%ev23, bem%ev31, bem%ev32, bem%ev33)
    deallocate(bem%yel, bem%yer, bem%xs4, bem%zs4, bem%ds)
  end subroutine hbiref_entries

  ! free-surface correction strains of one angular dislocation
  ! (AngDisStrainFSC, TDstressFS.f90:629-1011): out = v11 v22 v33 v12 v13 v23
  subroutine hbiref_angfsc(y1, y2, y3, beta, b1, b2, b3, nu, a, out) bind(C, name="hbiref_angfsc")
    real(c_double), value :: y1, y2, y3, beta, b1, b2, b3, nu, a
    real(c_double), intent(out) :: out(6)
    call AngDisStrainFSC(y1, y2, y3, beta, b1, b2, b3, nu, a, out(1), out(2), out(3), out(4), out(5), out(6))
  end subroutine hbiref_angfsc

  ! half-space stress of a triangular dislocation and its parts
  ! (TDstressHS, TDstressFS.f90:7-44; TDstress_HarFunc 236-293):
  ! out(1:6) total, out(7:12) harmonic part
  subroutine hbiref_td_hs(x, y, z, p1, p2, p3, ss, ds, ts, mu, lambda, out) bind(C, name="hbiref_td_hs")
    real(c_double), value :: x, y, z, ss, ds, ts, mu, lambda
    real(c_double), intent(in) :: p1(3), p2(3), p3(3)
    real(c_double), intent(out) :: out(12)
    call TDstressHS(x, y, z, p1, p2, p3, ss, ds, ts, mu, lambda, out(1), out(2), out(3), out(4), out(5), out(6))
    call TDstress_HarFunc(x, y, z, p1, p2, p3, ss, ds, ts, mu, lambda, out(7), out(8), out(9), out(10), out(11), &
      & out(12))
  end subroutine hbiref_td_hs
end module hbiref_harness
