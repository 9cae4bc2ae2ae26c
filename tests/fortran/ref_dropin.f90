! The reference's module globals and the drop-in (shared by drive_ref and
! drive_nve): a stand-in for the module globals the three-body routine reads
! and writes (common_arrays.f90:20-22 nom/nsite, :36-37 nm/ns, :43
! rcut3/rcutsd3, :46 alpha3/delta33, :100 fxs.., :110 vir, :201-218 the 3-body
! tables and COM-force/torque arrays), and the two snippets a maintainer adds
! (context creation; the new body of threebody_term), copied from
! the integration guide line for line (tests/test_fortran_binding.py checks that).
module ref_globals
  use, intrinsic :: iso_c_binding
  implicit none
  integer, parameter :: nom = 5000, nsite = 50
  integer :: nm, ns
  real(8) :: rcut3, rcutsd3, alpha3, delta33, vir
  logical :: S2_term, S3_term
  real(8), target :: bS3_s(nsite, nsite), gS3_s(nsite, nsite), r0S3_s(nsite, nsite)
  real(8), target :: bS3_l(nsite, nsite), gS3_l(nsite, nsite), r0S3_l(nsite, nsite)
  real(8), target :: cs3_s(5000), cs3_l(5000)
  real(8) :: fxs(nsite, nom), fys(nsite, nom), fzs(nsite, nom)
  real(8) :: fxfs2(nom), fyfs2(nom), fzfs2(nom), txfs2(nom), tyfs2(nom), tzfs2(nom)
  real(8) :: fxfs3(nom), fyfs3(nom), fzfs3(nom), txfs3(nom), tyfs3(nom), tzfs3(nom)
end module ref_globals

module threebody_dropin
  use, intrinsic :: iso_c_binding
  use ccpol3b_f
  use ref_globals
  implicit none
  type(c_ptr), save :: tb_ctx
  type(ccpol3b_params), target, save :: tb_par
contains

  subroutine create_context()
    integer(c_int) :: rc
    ! ---- integration snippet 1 (context creation), verbatim ----
    tb_par%beta_sr  = c_loc(bS3_s);  tb_par%gamma_sr = c_loc(gS3_s);  tb_par%r0_sr = c_loc(r0S3_s)
    tb_par%beta_lr  = c_loc(bS3_l);  tb_par%gamma_lr = c_loc(gS3_l);  tb_par%r0_lr = c_loc(r0S3_l)
    tb_par%ld_pair  = nsite
    tb_par%cs_sr    = c_loc(cs3_s);  tb_par%cs_lr = c_loc(cs3_l);     tb_par%nlin = 364
    tb_par%alpha3   = alpha3;  tb_par%delta33 = delta33;  tb_par%rcut3 = rcut3
    tb_par%s2_term  = merge(1, 0, S2_term);  tb_par%s3_term = merge(1, 0, S3_term)
    tb_par%precision = CCPOL3B_FP64          ! or CCPOL3B_MIXED
    tb_par%device    = -1                    ! current HIP device
    rc = ccpol3b_create(tb_par, tb_ctx)
    if (rc /= 0) stop 'ccpol3b_create failed'
    ! ---- end of snippet 1 ----
  end subroutine create_context

  ! ---- integration snippet 2 (the new body of threebody_term), verbatim ----
subroutine threebody_term(xxx,yyy,zzz,xxxs,yyys,zzzs,boxii,qfs2,qfs3)
  real(8) :: xxx(nom),yyy(nom),zzz(nom), xxxs(nsite,nom),yyys(nsite,nom),zzzs(nsite,nom)
  real(8) :: boxii, qfs2, qfs3
  integer :: ierr
  call ccpol3b_threebody_term_f(tb_ctx, nm, nsite, xxx, yyy, zzz, xxxs, yyys, zzzs, boxii, qfs2, qfs3, &
       fxs, fys, fzs, fxfs2, fyfs2, fzfs2, fxfs3, fyfs3, fzfs3, &
       txfs2, tyfs2, tzfs2, txfs3, tyfs3, tzfs3, vir, ierr)
  if (ierr /= 0) stop 'ccpol3b_threebody_term failed'
end subroutine threebody_term
  ! ---- end of snippet 2 ----

end module threebody_dropin

