! A host MD loop in Fortran that calls the drop-in threebody_term (ref_dropin.f90,
! the documented integration snippets) for its forces at every step, as the reference's
! NVE driver calls the original (evolve.f90:877): rigid waters, velocity Verlet
! for the centres of mass and, for the rotation, the symplectic splitting of
! the free rigid rotor of Dullweber, Leimkuhler and McLachlan (J. Chem. Phys.
! 107, 5840 (1997)): exact rotations about the body axes z, y, x, y, z for
! h/2, h/2, h, h/2, h/2, each turning the orientation quaternion and the
! body-frame angular momentum.  Units of the reference: A, ps, amu, 10 J/mol.
!
!   drive_nve IN OUT
! IN (text): nm box nsteps every dt(ps) / alpha3 delta33 rcut3 s2 s3 / SR path /
!            LR path / mass pmi1 pmi2 pmi3 / 17 lines of body-frame site
!            coordinates (A) / nm lines "x y z q0 q1 q2 q3 vx vy vz lx ly lz"
!            (COM A, orientation, COM velocity A/ps, body angular momentum
!            amu A^2/ps)
! OUT (text): per step "E step e3 kt kr" (kcal/mol); at step 0 and every
!            `every` steps "S step", then qfs2 qfs3 vir, nm lines x y z,
!            nm*17 lines xs ys zs (the site offsets passed to the call),
!            nm*17 lines fxs fys fzs, nm lines fxfs2.. fxfs3..
module rigid_water
  implicit none
contains

  ! rotation matrix (body -> lab) of the unit quaternion q = (q0, q1, q2, q3)
  pure function rotmat(q) result(a)
    real(8), intent(in) :: q(0:3)
    real(8) :: a(3, 3)
    a(1, 1) = q(0)**2 + q(1)**2 - q(2)**2 - q(3)**2
    a(1, 2) = 2.d0 * (q(1) * q(2) - q(0) * q(3))
    a(1, 3) = 2.d0 * (q(1) * q(3) + q(0) * q(2))
    a(2, 1) = 2.d0 * (q(1) * q(2) + q(0) * q(3))
    a(2, 2) = q(0)**2 - q(1)**2 + q(2)**2 - q(3)**2
    a(2, 3) = 2.d0 * (q(2) * q(3) - q(0) * q(1))
    a(3, 1) = 2.d0 * (q(1) * q(3) - q(0) * q(2))
    a(3, 2) = 2.d0 * (q(2) * q(3) + q(0) * q(1))
    a(3, 3) = q(0)**2 - q(1)**2 - q(2)**2 + q(3)**2
  end function rotmat

  ! rotation of the body by angle th about its own axis k: q <- q (x) (cos th/2, sin th/2 e_k),
  ! and the body-frame components of the (lab-fixed) angular momentum turn by -th
  subroutine body_turn(q, l, k, th)
    real(8), intent(inout) :: q(0:3), l(3)
    integer, intent(in) :: k
    real(8), intent(in) :: th
    real(8) :: c, s, r(0:3), p(0:3), u, w
    integer :: k1, k2
    c = cos(0.5d0 * th)
    s = sin(0.5d0 * th)
    r = 0.d0
    r(0) = c
    r(k) = s
    p(0) = q(0) * r(0) - q(1) * r(1) - q(2) * r(2) - q(3) * r(3)
    p(1) = q(0) * r(1) + q(1) * r(0) + q(2) * r(3) - q(3) * r(2)
    p(2) = q(0) * r(2) + q(2) * r(0) + q(3) * r(1) - q(1) * r(3)
    p(3) = q(0) * r(3) + q(3) * r(0) + q(1) * r(2) - q(2) * r(1)
    q = p
    k1 = mod(k, 3) + 1
    k2 = mod(k + 1, 3) + 1
    c = cos(th)
    s = sin(th)
    u = l(k1)
    w = l(k2)
    l(k1) = c * u + s * w
    l(k2) = -s * u + c * w
  end subroutine body_turn

  ! free rotation over h (angular velocity about body axis k: l(k) / pmi(k))
  subroutine free_rotor(q, l, pmi, h)
    real(8), intent(inout) :: q(0:3), l(3)
    real(8), intent(in) :: pmi(3), h
    call body_turn(q, l, 3, 0.5d0 * h * l(3) / pmi(3))
    call body_turn(q, l, 2, 0.5d0 * h * l(2) / pmi(2))
    call body_turn(q, l, 1, h * l(1) / pmi(1))
    call body_turn(q, l, 2, 0.5d0 * h * l(2) / pmi(2))
    call body_turn(q, l, 3, 0.5d0 * h * l(3) / pmi(3))
    q = q / sqrt(sum(q * q))
  end subroutine free_rotor
end module rigid_water

program drive_nve
  use, intrinsic :: iso_c_binding
  use ccpol3b_f
  use ref_globals
  use threebody_dropin
  use rigid_water
  implicit none
  character(len=4096) :: fin, fout, psr, plr
  integer :: i, s, is2, is3, rc, nsteps, every, step
  integer(c_int) :: nl
  real(8) :: boxl, qfs2, qfs3, dt, mass, pmi(3), pc(3, 17), a(3, 3), f(3), t(3), kt, kr
  real(8), allocatable :: x(:), y(:), z(:), xs(:, :), ys(:, :), zs(:, :)
  real(8), allocatable :: q(:, :), v(:, :), l(:, :)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open (10, file=trim(fin), status='old', action='read')
  read (10, *) nm, boxl, nsteps, every, dt
  read (10, *) alpha3, delta33, rcut3, is2, is3
  read (10, '(a)') psr
  read (10, '(a)') plr
  read (10, *) mass, pmi
  do s = 1, 17
    read (10, *) pc(:, s)
  end do
  ns = 28
  S2_term = is2 /= 0
  S3_term = is3 /= 0
  rcutsd3 = rcut3**2
  nl = 5000
  rc = ccpol3b_read_data3b(trim(psr)//c_null_char, bS3_s, gS3_s, r0S3_s, int(nsite, c_int), cs3_s, nl)
  if (rc /= 0 .or. nl /= 364) stop 'read SR'
  nl = 5000
  rc = ccpol3b_read_data3b(trim(plr)//c_null_char, bS3_l, gS3_l, r0S3_l, int(nsite, c_int), cs3_l, nl)
  if (rc /= 0 .or. nl /= 364) stop 'read LR'
  allocate (x(nom), y(nom), z(nom), xs(nsite, nom), ys(nsite, nom), zs(nsite, nom))
  allocate (q(0:3, nm), v(3, nm), l(3, nm))
  x = 0.d0; y = 0.d0; z = 0.d0; xs = 0.d0; ys = 0.d0; zs = 0.d0
  do i = 1, nm
    read (10, *) x(i), y(i), z(i), q(:, i), v(:, i), l(:, i)
  end do
  close (10)

  call create_context()
  open (11, file=trim(fout), status='replace', action='write')
  call forces()
  call record(0)
  do step = 1, nsteps
    call kick()
    do i = 1, nm
      x(i) = x(i) + dt * v(1, i)
      y(i) = y(i) + dt * v(2, i)
      z(i) = z(i) + dt * v(3, i)
      call free_rotor(q(:, i), l(:, i), pmi, dt)
    end do
    call forces()
    call kick()
    call record(step)
  end do
  close (11)
  call ccpol3b_destroy(tb_ctx)
  print '(a)', 'ok'

contains

  ! site offsets from the orientations, then the three-body call; site forces
  ! and the virial are zeroed first, as the two-body routine leaves them for it
  subroutine forces()
    integer :: m, js
    do m = 1, nm
      a = rotmat(q(:, m))
      do js = 1, 17
        xs(js, m) = a(1, 1) * pc(1, js) + a(1, 2) * pc(2, js) + a(1, 3) * pc(3, js)
        ys(js, m) = a(2, 1) * pc(1, js) + a(2, 2) * pc(2, js) + a(2, 3) * pc(3, js)
        zs(js, m) = a(3, 1) * pc(1, js) + a(3, 2) * pc(2, js) + a(3, 3) * pc(3, js)
      end do
    end do
    fxs = 0.d0; fys = 0.d0; fzs = 0.d0; vir = 0.d0
    call threebody_term(x, y, z, xs, ys, zs, boxl, qfs2, qfs3)
  end subroutine forces

  ! half step of the velocities and the body angular momenta: force = site
  ! forces + COM forces, torque = sum of site offset x site force (lab), turned
  ! into the body frame
  subroutine kick()
    integer :: m, js
    do m = 1, nm
      f = 0.d0
      t = 0.d0
      do js = 1, 17
        f(1) = f(1) + fxs(js, m)
        f(2) = f(2) + fys(js, m)
        f(3) = f(3) + fzs(js, m)
        t(1) = t(1) + ys(js, m) * fzs(js, m) - zs(js, m) * fys(js, m)
        t(2) = t(2) + zs(js, m) * fxs(js, m) - xs(js, m) * fzs(js, m)
        t(3) = t(3) + xs(js, m) * fys(js, m) - ys(js, m) * fxs(js, m)
      end do
      f(1) = f(1) + fxfs2(m) + fxfs3(m)
      f(2) = f(2) + fyfs2(m) + fyfs3(m)
      f(3) = f(3) + fzfs2(m) + fzfs3(m)
      v(:, m) = v(:, m) + 0.5d0 * dt * f / mass
      a = rotmat(q(:, m))
      l(:, m) = l(:, m) + 0.5d0 * dt * matmul(transpose(a), t)
    end do
  end subroutine kick

  subroutine record(st)
    integer, intent(in) :: st
    integer :: m, js
    kt = 0.d0
    kr = 0.d0
    do m = 1, nm
      kt = kt + 0.5d0 * mass * sum(v(:, m)**2)
      kr = kr + 0.5d0 * sum(l(:, m)**2 / pmi)
    end do
    write (11, '(a,i8,3es26.17)') 'E', st, qfs2 + qfs3, kt / 418.4d0, kr / 418.4d0
    if (mod(st, every) /= 0) return
    write (11, '(a,i8)') 'S', st
    write (11, '(3es26.17)') qfs2, qfs3, vir
    do m = 1, nm
      write (11, '(3es26.17)') x(m), y(m), z(m)
    end do
    do m = 1, nm
      do js = 1, 17
        write (11, '(3es26.17)') xs(js, m), ys(js, m), zs(js, m)
      end do
    end do
    do m = 1, nm
      do js = 1, 17
        write (11, '(3es26.17)') fxs(js, m), fys(js, m), fzs(js, m)
      end do
    end do
    do m = 1, nm
      write (11, '(6es26.17)') fxfs2(m), fyfs2(m), fzfs2(m), fxfs3(m), fyfs3(m), fzfs3(m)
    end do
  end subroutine record
end program drive_nve
