! ccpol3b_f.f90 -- Fortran 2003 (iso_c_binding) interface to libccpol3b.so.
!
! The C entry points are declared in ccpol3b.h; this module binds them for a
! Fortran caller and adds ccpol3b_threebody_term_f, which has the argument list
! and side effects of the reference routine
!     threebody_term(xxx,yyy,zzz,xxxs,yyys,zzzs,boxii,qfs2,qfs3)
! (source/threebody_potential.f90:17-99) with the module globals it reads and
! writes passed explicitly: site forces fxs,fys,fzs(ld,nm) are accumulated,
! COM forces fxfs2..,fxfs3.. overwritten, torques zeroed, vir accumulated.
module ccpol3b_f
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: CCPOL3B_FP64 = 0, CCPOL3B_MIXED = 1
  integer(c_int), parameter, public :: CCPOL3B_OK = 0, CCPOL3B_EINVAL = -1, CCPOL3B_EHIP = -2, &
                                       CCPOL3B_ENOMEM = -3, CCPOL3B_ECAPACITY = -4, CCPOL3B_ENODEV = -5

  ! struct ccpol3b_params (ccpol3b.h)
  type, bind(C), public :: ccpol3b_params
    type(c_ptr) :: beta_sr = c_null_ptr, gamma_sr = c_null_ptr, r0_sr = c_null_ptr
    type(c_ptr) :: beta_lr = c_null_ptr, gamma_lr = c_null_ptr, r0_lr = c_null_ptr
    integer(c_int) :: ld_pair = 0
    type(c_ptr) :: cs_sr = c_null_ptr, cs_lr = c_null_ptr
    integer(c_int) :: nlin = 364
    real(c_double) :: alpha3 = 0.d0, delta33 = 0.d0, rcut3 = 0.d0
    integer(c_int) :: s2_term = 1, s3_term = 1
    integer(c_int) :: precision = CCPOL3B_FP64
    integer(c_int) :: device = -1
  end type

  ! struct ccpol3b_stats (ccpol3b.h)
  type, bind(C), public :: ccpol3b_stats
    integer(c_long) :: n_fs3, n_fs3_sr, n_fs3_switch, n_fs3_lr, n_fs2_items, n_fs2_pairs
    real(c_double) :: ms_total, ms_lists, ms_fs2, ms_fs3, ms_reduce, ms_h2d, ms_d2h, ms_fs3_eval
  end type

  public :: ccpol3b_create, ccpol3b_destroy, ccpol3b_set_precision, ccpol3b_threebody_term, &
            ccpol3b_last_stats, ccpol3b_read_data3b, ccpol3b_orbit_index, ccpol3b_threebody_term_f, &
            ccpol3b_md_create, ccpol3b_md_set_state, ccpol3b_md_get_state, ccpol3b_md_run, ccpol3b_md_destroy

  interface
    integer(c_int) function ccpol3b_create(params, ctx) bind(C, name="ccpol3b_create")
      import :: c_int, c_ptr, ccpol3b_params
      type(ccpol3b_params), intent(in) :: params
      type(c_ptr), intent(out) :: ctx
    end function

    subroutine ccpol3b_destroy(ctx) bind(C, name="ccpol3b_destroy")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine

    integer(c_int) function ccpol3b_set_precision(ctx, precision) bind(C, name="ccpol3b_set_precision")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: precision
    end function

    integer(c_int) function ccpol3b_threebody_term(ctx, nm, ld_site, x, y, z, xs, ys, zs, box, qfs2, qfs3, &
        fxs, fys, fzs, fxfs2, fyfs2, fzfs2, fxfs3, fyfs3, fzfs3, txfs2, tyfs2, tzfs2, txfs3, tyfs3, tzfs3, vir) &
        bind(C, name="ccpol3b_threebody_term")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nm, ld_site
      real(c_double), intent(in) :: x(*), y(*), z(*), xs(*), ys(*), zs(*)
      real(c_double), value :: box
      real(c_double), intent(out) :: qfs2, qfs3
      real(c_double), intent(inout) :: fxs(*), fys(*), fzs(*)
      real(c_double), intent(inout) :: fxfs2(*), fyfs2(*), fzfs2(*), fxfs3(*), fyfs3(*), fzfs3(*)
      real(c_double), intent(inout) :: txfs2(*), tyfs2(*), tzfs2(*), txfs3(*), tyfs3(*), tzfs3(*)
      real(c_double), intent(inout) :: vir
    end function

    integer(c_int) function ccpol3b_last_stats(ctx, stats) bind(C, name="ccpol3b_last_stats")
      import :: c_int, c_ptr, ccpol3b_stats
      type(c_ptr), value :: ctx
      type(ccpol3b_stats), intent(out) :: stats
    end function

    integer(c_int) function ccpol3b_read_data3b(path, beta, gamma, r0, ld_pair, cs, nlin) &
        bind(C, name="ccpol3b_read_data3b")
      import :: c_int, c_double, c_char
      character(kind=c_char), intent(in) :: path(*)
      real(c_double), intent(out) :: beta(*), gamma(*), r0(*), cs(*)
      integer(c_int), value :: ld_pair
      integer(c_int), intent(inout) :: nlin
    end function

    integer(c_int) function ccpol3b_orbit_index(is, js, ks, ip, jp, kp) bind(C, name="ccpol3b_orbit_index")
      import :: c_int
      integer(c_int), value :: is, js, ks, ip, jp, kp
    end function

    ! Rigid-water NVE driver on the device (velocity Verlet + NO_SQUISH, the
    ! reference's nve_1, evolve.f90:621-1062, 8107-8204).  Fortran arrays map
    ! onto the C layouts directly: pcoord(3,ns), com/vel/jbody(3,nm),
    ! quat(4,nm), records(4,nsteps).
    integer(c_int) function ccpol3b_md_create(ctx, nm, box, ns, pcoord, mass, pmi, md) &
        bind(C, name="ccpol3b_md_create")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nm, ns
      real(c_double), value :: box, mass
      real(c_double), intent(in) :: pcoord(*), pmi(3)
      type(c_ptr), intent(out) :: md
    end function

    integer(c_int) function ccpol3b_md_set_state(md, com, quat, vel, jbody) bind(C, name="ccpol3b_md_set_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: md
      real(c_double), intent(in) :: com(*), quat(*), vel(*), jbody(*)
    end function

    integer(c_int) function ccpol3b_md_get_state(md, com, quat, vel, jbody) bind(C, name="ccpol3b_md_get_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: md
      real(c_double), intent(out) :: com(*), quat(*), vel(*), jbody(*)
    end function

    integer(c_int) function ccpol3b_md_run(md, nsteps, dt, records) bind(C, name="ccpol3b_md_run")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: md
      integer(c_int), value :: nsteps
      real(c_double), value :: dt
      real(c_double), intent(out) :: records(*)
    end function

    subroutine ccpol3b_md_destroy(md) bind(C, name="ccpol3b_md_destroy")
      import :: c_ptr
      type(c_ptr), value :: md
    end subroutine
  end interface

contains

  ! threebody_term with explicit globals.  ld is the site leading dimension of
  ! xxxs..(ld,nm) and fxs..(ld,nm) (nsite = 50 in common_arrays.f90:22).
  ! Returns the library status in ierr (0 on success; outputs untouched otherwise).
  subroutine ccpol3b_threebody_term_f(ctx, nm, ld, xxx, yyy, zzz, xxxs, yyys, zzzs, boxii, qfs2, qfs3, &
      fxs, fys, fzs, fxfs2, fyfs2, fzfs2, fxfs3, fyfs3, fzfs3, txfs2, tyfs2, tzfs2, txfs3, tyfs3, tzfs3, &
      vir, ierr)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: nm, ld
    real(8), intent(in) :: xxx(nm), yyy(nm), zzz(nm)
    real(8), intent(in) :: xxxs(ld, nm), yyys(ld, nm), zzzs(ld, nm)
    real(8), intent(in) :: boxii
    real(8), intent(out) :: qfs2, qfs3
    real(8), intent(inout) :: fxs(ld, nm), fys(ld, nm), fzs(ld, nm)
    real(8), intent(inout) :: fxfs2(nm), fyfs2(nm), fzfs2(nm), fxfs3(nm), fyfs3(nm), fzfs3(nm)
    real(8), intent(inout) :: txfs2(nm), tyfs2(nm), tzfs2(nm), txfs3(nm), tyfs3(nm), tzfs3(nm)
    real(8), intent(inout) :: vir
    integer, intent(out) :: ierr
    ierr = ccpol3b_threebody_term(ctx, int(nm, c_int), int(ld, c_int), xxx, yyy, zzz, xxxs, yyys, zzzs, &
        real(boxii, c_double), qfs2, qfs3, fxs, fys, fzs, fxfs2, fyfs2, fzfs2, fxfs3, fyfs3, fzfs3, &
        txfs2, tyfs2, tzfs2, txfs3, tyfs3, tzfs3, vir)
  end subroutine

end module ccpol3b_f
