! The reference's own call path with libccpol3b.so dropped in, at the reference's
! array shapes (the globals and the drop-in of ref_dropin.f90), the parameter
! reads of initialization.f90:640-701 (here through ccpol3b_read_data3b into
! the (nsite,nsite) tables, ld_pair = 50) and rcutsd3 = rcut3**2 as
! initialization.f90:158 sets it.
!
!   drive_ref IN OUT
! IN (text): nm box / alpha3 delta33 rcut3 s2 s3 / SR path / LR path /
!            nm lines "x y z" (A) / nm*17 lines "xs ys zs" (A, site-major per molecule)
! OUT (text): qfs2 qfs3 vir / nm*17 lines fxs fys fzs / nm lines fxfs2.. fxfs3..
program drive_ref
  use, intrinsic :: iso_c_binding
  use ccpol3b_f
  use ref_globals
  use threebody_dropin
  implicit none
  character(len=4096) :: fin, fout, psr, plr
  integer :: i, s, is2, is3, rc
  integer(c_int) :: nl
  real(8) :: boxl, qfs2, qfs3
  real(8), allocatable :: x(:), y(:), z(:), xs(:, :), ys(:, :), zs(:, :)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open (10, file=trim(fin), status='old', action='read')
  read (10, *) nm, boxl
  read (10, *) alpha3, delta33, rcut3, is2, is3
  read (10, '(a)') psr
  read (10, '(a)') plr
  ns = 28
  S2_term = is2 /= 0
  S3_term = is3 /= 0
  rcutsd3 = rcut3**2   ! initialization.f90:158; the library takes rcut3 itself
  ! the parameter reads of initialization.f90:660-701 into the (nsite,nsite) tables
  nl = 5000
  rc = ccpol3b_read_data3b(trim(psr)//c_null_char, bS3_s, gS3_s, r0S3_s, int(nsite, c_int), cs3_s, nl)
  if (rc /= 0 .or. nl /= 364) stop 'read SR'
  nl = 5000
  rc = ccpol3b_read_data3b(trim(plr)//c_null_char, bS3_l, gS3_l, r0S3_l, int(nsite, c_int), cs3_l, nl)
  if (rc /= 0 .or. nl /= 364) stop 'read LR'

  allocate (x(nom), y(nom), z(nom), xs(nsite, nom), ys(nsite, nom), zs(nsite, nom))
  x = 0.d0; y = 0.d0; z = 0.d0; xs = 0.d0; ys = 0.d0; zs = 0.d0
  do i = 1, nm
    read (10, *) x(i), y(i), z(i)
  end do
  do i = 1, nm
    do s = 1, 17
      read (10, *) xs(s, i), ys(s, i), zs(s, i)
    end do
  end do
  close (10)

  call create_context()
  ! what the caller has done before the call (two-body routine / zero_site_forces):
  ! site forces and virial zeroed; COM forces and torques hold stale values
  fxs = 0.d0; fys = 0.d0; fzs = 0.d0; vir = 0.d0
  fxfs2 = 7.d0; fyfs2 = 7.d0; fzfs2 = 7.d0; fxfs3 = 7.d0; fyfs3 = 7.d0; fzfs3 = 7.d0
  txfs2 = 7.d0; tyfs2 = 7.d0; tzfs2 = 7.d0; txfs3 = 7.d0; tyfs3 = 7.d0; tzfs3 = 7.d0
  qfs2 = 0.d0; qfs3 = 0.d0
  call threebody_term(x, y, z, xs, ys, zs, boxl, qfs2, qfs3)
  call ccpol3b_destroy(tb_ctx)
  if (S2_term .and. maxval(abs(txfs2(1:nm))) /= 0.d0) stop 'FS2 torques not zeroed'
  if (S3_term .and. maxval(abs(txfs3(1:nm))) /= 0.d0) stop 'FS3 torques not zeroed'
  if (.not. S2_term .and. minval(fxfs2(1:nm)) /= 7.d0) stop 'FS2 COM forces touched'
  if (maxval(abs(fxs(18:nsite, 1:nm))) /= 0.d0) stop 'sites beyond 17 touched'
  if (maxval(abs(fxs(:, nm + 1:nom))) /= 0.d0) stop 'molecules beyond nm touched'
  open (11, file=trim(fout), status='replace', action='write')
  write (11, '(3es26.17)') qfs2, qfs3, vir
  do i = 1, nm
    do s = 1, 17
      write (11, '(3es26.17)') fxs(s, i), fys(s, i), fzs(s, i)
    end do
  end do
  do i = 1, nm
    write (11, '(6es26.17)') fxfs2(i), fyfs2(i), fzfs2(i), fxfs3(i), fyfs3(i), fzfs3(i)
  end do
  close (11)
  print '(a)', 'ok'
end program drive_ref
