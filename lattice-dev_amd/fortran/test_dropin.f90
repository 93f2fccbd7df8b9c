!
! Exercises lattice_dropin the way a reference driver calls its modules
! (Pseudospec/openmp_dev/evolve_spectral.f90: laplacian_3d, init_spectrum_3d,
! spectrum_3d, calc_metric, symp4).  Prints DROPIN OK or stops with code 1.
!
program test_dropin
  use, intrinsic :: iso_c_binding
  use lattice_dropin
  implicit none
  integer, parameter :: dl = kind(1.d0)
  integer, parameter :: n = 16, nfld = 2
  real(dl), parameter :: len = 10._dl, twopi = 6.2831853071795864769252867665590_dl
  real(dl) :: dk, dx, err, x, y, z, kk2, ysc, yscp, e(LAT_NE)
  real(c_double), allocatable, target :: f(:,:,:), g(:,:,:)
  complex(c_double_complex), allocatable, target :: Fk(:,:,:)
  real(dl), allocatable :: S(:)
  real(dl), allocatable :: fld(:,:,:,:), fldp(:,:,:,:)
  type(c_ptr) :: planf, planb
  integer :: i, j, k, b

  dk = twopi / len
  dx = len / n
  call lattice_init(n, n, n, nfld, len, LAT_F64, LAT_LAP_SPECTRAL, LAT_COSMIC, LAT_MODEL_TWOFIELD, &
                    (/ 1._dl, 2.5e5_dl /))

  allocate(f(n,n,n), g(n,n,n), Fk(n/2+1,n,n))
  kk2 = dk**2 * (3**2 + 2**2 + 1**2)
  do k = 1, n; do j = 1, n; do i = 1, n
     x = (i-1)*dx; y = (j-1)*dx; z = (k-1)*dx
     f(i,j,k) = cos(dk*(3*x + 2*y - z))
  end do; end do; end do
  g = f
  planf = c_null_ptr; planb = c_null_ptr
  call laplacian_3d(n, n, n, g, Fk, dk, planf, planb)
  err = maxval(abs(g + kk2*f))
  print *, "laplacian_3d plane wave max error", err
  if (err > 1e-10_dl) stop 1

  call init_spectrum_3d(n, n, n)
  allocate(S(ns))
  g = f
  call spectrum_3d(S, g, Fk)
  ! |k| = sqrt(14) = 3.74: weight only in bins 3 and 4 (1-based 4, 5)
  do b = 1, ns
     if ((b == 4 .or. b == 5) .neqv. (abs(S(b)) > 1e-8_dl * maxval(abs(S)))) then
        print *, "spectrum bin", b, S(b)
        stop 1
     end if
  end do
  print *, "spectrum_3d bins", ns, " peak", maxval(S)

  ! Discrete-tree spectra of the same plane wave (Discrete/Entropy/analysis.f90)
  deallocate(S)
  allocate(S(int(sqrt(3._dl) * (n / 2) + 2)))
  g = f
  call spectrum_bin(g, S)
  do b = 1, size(S)
     if ((b == 4) .neqv. (abs(S(b)) > 1e-8_dl * maxval(abs(S)))) then
        print *, "spectrum_bin bin", b, S(b)
        stop 1
     end if
  end do
  g = f
  call spectrum(g, S)
  if (abs(S(4)) + abs(S(5)) < 1e-8_dl .or. abs(S(3)) > 1e-8_dl * maxval(abs(S))) stop 1
  print *, "Discrete spectrum / spectrum_bin ok"
  deallocate(S)
  allocate(S(ns))

  allocate(fld(nfld,0:n+1,0:n+1,0:n+1), fldp(nfld,n,n,n))
  fld = 0._dl; fldp = 0._dl
  do k = 1, n; do j = 1, n; do i = 1, n
     fld(1,i,j,k) = 1.0093430384226378929425913902459_dl + 1e-6_dl*f(i,j,k)
     fld(2,i,j,k) = 1e-6_dl*f(j,k,i)
     fldp(1,i,j,k) = -0.7137133070120812430962278466136_dl
  end do; end do; end do
  call lattice_put_state(fld, fldp, 1, 1._dl, 0._dl)
  call calc_metric()
  call symp4(1e-3_dl, 5)
  call lattice_get_state(fld, fldp, 1, ysc, yscp)
  call lattice_energies(e)
  print *, "after symp4: a", get_scale_factor(), " H", get_hubble(), " resid", e(8)
  if (abs(e(8)) > 1e-9_dl) stop 1
  ! O(1) background reads agree with the reduction-based diagnostics
  if (abs(get_scale_factor() - e(1)) > 1e-15_dl * e(1)) stop 1
  if (abs(get_hubble() - e(2)) > 1e-15_dl * abs(e(2))) stop 1
  if (abs(grav_energy() - e(7)) > 1e-15_dl * abs(e(7))) stop 1
  if (abs(kinetic_norm() - 1._dl / ysc**4) > 1e-15_dl / ysc**4) stop 1
  ! mpi_dev dump_rho sequence (mpi_dev/evolve_spectral.f90:304-315)
  block
    real(dl) :: rho, ke, ge, pe
    rho = scalar_rho()
    ke = kinetic_energy(); ge = grad_energy(); pe = potential_energy()
    print *, "scalar_rho", rho, " KE", ke, " GE", ge, " PE", pe, " resid", (rho + grav_energy()) / rho
    if (abs(rho - e(3)) > 1e-14_dl * e(3)) stop 1
    if (abs(ke + pe + ge - rho) > 1e-14_dl * rho) stop 1
    if (abs(ke - e(4)) > 1e-14_dl * e(4) .or. abs(pe - e(5)) > 1e-14_dl * e(5)) stop 1
    if (abs((rho + grav_energy()) / rho - e(8)) > 1e-12_dl) stop 1
    if (grad_en() == grad_en()) stop 1   ! spectral context: NaN
  end block
  if (abs(fld(1,0,5,5) - fld(1,n,5,5)) > 0._dl) stop 1   ! periodic ghost
  if (.not. (ysc > 1._dl)) stop 1

  ! Discrete-tree gradient energies on a 7-point FD context: GRAD2 form ==
  ! Laplacian form (summation by parts, homlat.f90:88-147)
  call lattice_init(n, n, n, nfld, len, LAT_F64, LAT_LAP_STENCIL, LAT_COSMIC, LAT_MODEL_TWOFIELD, &
                    (/ 1._dl, 2.5e5_dl /))
  do k = 1, n; do j = 1, n; do i = 1, n
     fld(1,i,j,k) = 1._dl + 0.01_dl*f(i,j,k)
     fld(2,i,j,k) = 0.01_dl*f(j,k,i)
  end do; end do; end do
  fldp = 0._dl
  call lattice_put_state(fld, fldp, 1, 1._dl, 0._dl)
  print *, "grad_en", grad_en(), " grad_en_wlap", grad_en_wlap()
  if (.not. (grad_en() > 0._dl)) stop 1
  if (abs(grad_en() - grad_en_wlap()) > 1e-12_dl * grad_en_wlap()) stop 1
  ! the Discrete tree's make_output writes LOG.out and PSD in the working directory
  call discrete_make_output(0._dl, 1e-3_dl, 2e5_dl, 1)
  block
    logical :: ok1, ok2
    inquire(file="LOG.out", exist=ok1)
    inquire(file="PSD", exist=ok2)
    if (.not. (ok1 .and. ok2)) stop 1
  end block

  ! metric perturbations sourced by a two-field configuration with gradients
  call lattice_init(n, n, n, nfld, len, LAT_F64, LAT_LAP_SPECTRAL, LAT_COSMIC, LAT_MODEL_TWOFIELD, &
                    (/ 1._dl, 100._dl /), gw=.true., gw_mpl=1._dl)
  do k = 1, n; do j = 1, n; do i = 1, n
     x = (i-1)*dx; y = (j-1)*dx
     fld(1,i,j,k) = 1._dl + 0.01_dl*f(i,j,k)
     fld(2,i,j,k) = 0.01_dl*cos(dk*x)*cos(dk*y)   ! source with a transverse-traceless part
  end do; end do; end do
  fldp = 0._dl
  call lattice_put_state(fld, fldp, 1, 1._dl, 0._dl)
  block
    real(dl), allocatable :: hij(:,:,:,:,:), pij(:,:,:,:,:), Sp(:), Sh(:)
    allocate(hij(2,3,n,n,n), pij(2,3,n,n,n), Sp(ns), Sh(ns))
    hij = 0._dl; pij = 0._dl
    call lattice_put_gw(hij, pij)
    call calc_metric()
    call symp4(1e-2_dl, 5)
    call lattice_get_gw(hij, pij)
    call gw_spectrum(Sp, Sh)
    print *, "gw: max |pi_h|", maxval(abs(pij)), " peak S_pi", maxval(Sp)
    if (.not. (maxval(abs(pij)) > 1e-8_dl) .or. .not. (maxval(Sp) > 1e-12_dl)) stop 1
  end block
  call lattice_finalize()
  print *, "DROPIN OK"
end program test_dropin
