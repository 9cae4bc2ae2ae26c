! Fortran caller of libccpol3b.so through the ccpol3b_f module: the way the
! reference MD loop (evolve.f90) would call the three-body term.
!
!   drive_box probe            -- reads both parameter files named in the input,
!                                 prints orbit numbers and the status of
!                                 ccpol3b_create (no evaluation)
!   drive_box run IN OUT       -- evaluates the box in IN, writes OUT
!
! IN (text): nm ld box / alpha3 delta33 rcut3 s2 s3 precision / SR path / LR path /
!            nm lines "x y z" (A) / nm*17 lines "xs ys zs" (A, site-major per molecule)
! OUT (text): qfs2 qfs3 vir / nm*17 lines fxs fys fzs / nm lines fxfs2.. fxfs3..
program drive_box
  use, intrinsic :: iso_c_binding
  use ccpol3b_f
  implicit none
  character(len=4096) :: mode, fin, fout, psr, plr
  integer :: nm, ld, s2, s3, prec, i, s, ierr, rc
  real(8) :: box, alpha3, delta33, rcut3, qfs2, qfs3, vir
  real(c_double), target :: bsr(17, 17), gsr(17, 17), rsr(17, 17), blr(17, 17), glr(17, 17), rlr(17, 17)
  real(c_double), target :: csr(364), clr(364)
  integer(c_int) :: nl
  real(8), allocatable :: x(:), y(:), z(:), xs(:, :), ys(:, :), zs(:, :), fxs(:, :), fys(:, :), fzs(:, :)
  real(8), allocatable :: f2(:, :), f3(:, :), t2(:, :), t3(:, :)
  type(ccpol3b_params), target :: p
  type(c_ptr) :: ctx

  call get_command_argument(1, mode)
  call get_command_argument(2, fin)
  call get_command_argument(3, fout)
  open (10, file=trim(fin), status='old', action='read')
  read (10, *) nm, ld, box
  read (10, *) alpha3, delta33, rcut3, s2, s3, prec
  read (10, '(a)') psr
  read (10, '(a)') plr
  nl = 364
  rc = ccpol3b_read_data3b(trim(psr)//c_null_char, bsr, gsr, rsr, 17_c_int, csr, nl)
  if (rc /= 0) stop 'read SR'
  nl = 364
  rc = ccpol3b_read_data3b(trim(plr)//c_null_char, blr, glr, rlr, 17_c_int, clr, nl)
  if (rc /= 0) stop 'read LR'
  p%beta_sr = c_loc(bsr); p%gamma_sr = c_loc(gsr); p%r0_sr = c_loc(rsr)
  p%beta_lr = c_loc(blr); p%gamma_lr = c_loc(glr); p%r0_lr = c_loc(rlr)
  p%ld_pair = 17
  p%cs_sr = c_loc(csr); p%cs_lr = c_loc(clr)
  p%nlin = 364
  p%alpha3 = alpha3; p%delta33 = delta33; p%rcut3 = rcut3
  p%s2_term = s2; p%s3_term = s3; p%precision = prec; p%device = -1

  if (trim(mode) == 'probe') then
    print '(a,i0,1x,i0,1x,i0)', 'orbits ', ccpol3b_orbit_index(1, 1, 1, 0, 0, 0), &
        ccpol3b_orbit_index(17, 17, 17, 1, 1, 1), ccpol3b_orbit_index(2, 3, 1, 0, 1, 0)
    print '(a,es24.16,1x,es24.16)', 'coef ', csr(1), clr(364)
    rc = ccpol3b_create(p, ctx)
    print '(a,i0)', 'create ', rc
    if (rc == 0) call ccpol3b_destroy(ctx)
    stop
  end if

  allocate (x(nm), y(nm), z(nm), xs(ld, nm), ys(ld, nm), zs(ld, nm))
  allocate (fxs(ld, nm), fys(ld, nm), fzs(ld, nm), f2(nm, 3), f3(nm, 3), t2(nm, 3), t3(nm, 3))
  xs = 0.d0; ys = 0.d0; zs = 0.d0
  do i = 1, nm
    read (10, *) x(i), y(i), z(i)
  end do
  do i = 1, nm
    do s = 1, 17
      read (10, *) xs(s, i), ys(s, i), zs(s, i)
    end do
  end do
  close (10)
  rc = ccpol3b_create(p, ctx)
  if (rc /= 0) then
    print '(a,i0)', 'create failed ', rc
    stop 2
  end if
  fxs = 0.d0; fys = 0.d0; fzs = 0.d0; vir = 0.d0; f2 = 0.d0; f3 = 0.d0
  t2 = 1.d0; t3 = 1.d0
  call ccpol3b_threebody_term_f(ctx, nm, ld, x, y, z, xs, ys, zs, box, qfs2, qfs3, fxs, fys, fzs, &
      f2(:, 1), f2(:, 2), f2(:, 3), f3(:, 1), f3(:, 2), f3(:, 3), &
      t2(:, 1), t2(:, 2), t2(:, 3), t3(:, 1), t3(:, 2), t3(:, 3), vir, ierr)
  call ccpol3b_destroy(ctx)
  if (ierr /= 0) then
    print '(a,i0)', 'evaluation failed ', ierr
    stop 3
  end if
  ! torques of the active terms are zeroed; an inactive term's arrays are not touched
  if (s2 /= 0 .and. maxval(abs(t2)) /= 0.d0) stop 'FS2 torques not zeroed'
  if (s3 /= 0 .and. maxval(abs(t3)) /= 0.d0) stop 'FS3 torques not zeroed'
  if (s2 == 0 .and. minval(t2) /= 1.d0) stop 'FS2 torques touched'
  open (11, file=trim(fout), status='replace', action='write')
  write (11, '(3es26.17)') qfs2, qfs3, vir
  do i = 1, nm
    do s = 1, 17
      write (11, '(3es26.17)') fxs(s, i), fys(s, i), fzs(s, i)
    end do
  end do
  do i = 1, nm
    write (11, '(6es26.17)') f2(i, 1), f2(i, 2), f2(i, 3), f3(i, 1), f3(i, 2), f3(i, 3)
  end do
  close (11)
  print '(a)', 'ok'
end program drive_box
