!
! lattice_dropin -- Fortran drop-in for the reference's evolution-path
! procedures, over the C ABI of include/lattice_c.h (iso_c_binding).
!
! A reference driver keeps its call sites: laplacian_3d, init_spectrum_3d,
! spectrum_3d, crossspec_3d (Pseudospec/openmp_dev/fftw_mod.f90:177-207,
! analysis_spec.f90:48-178), Hamiltonian_Split, symp_o2step, calc_metric,
! get_scale_factor, get_hubble, kinetic_norm, grav_energy
! (hamiltonian_conformal.f90:50-304), the mpi_dev accessors scalar_rho,
! kinetic_energy, grad_energy, potential_energy
! (Pseudospec/mpi_dev/hamiltonian_cosmic.f90:255-349), the Discrete-tree
! grad_en / grad_en_wlap (Discrete/Entropy/homlat.f90:88-147) and symp2/4/6/8
! (Pseudospec/mpi_dev/integrator.f90:11-110), replacing `use fftw3`,
! `use analysis`, `use Hamiltonian`, `use integrator` by `use lattice_dropin`.
!
! Differences a caller must know:
!  * the state (fld, fldp, yscl, ysclp) lives on the GPU; lattice_put_state
!    uploads the reference's arrays fld(nfld,0:nx+1,0:ny+1,0:nz+1) /
!    fldp(nfld,nx,ny,nz), lattice_get_state downloads them (ghosts filled
!    periodically, as wrap_fields would);
!  * Fk and FFTW plan arguments are accepted and ignored (the context owns
!    its transforms), except that spectrum_3d returns the R2C in Fk as the
!    reference leaves it there;
!  * a nonzero status prints the library's message and stops, as the
!    reference does for an undefined split index.
!
module lattice_dropin
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer, parameter :: dl = kind(1.d0)

  type, bind(C) :: lat_config
     integer(c_int) :: nx, ny, nz, nfld
     real(c_double) :: len
     integer(c_int) :: precision, laplacian
     real(c_double) :: stencil(5)
     integer(c_int) :: time_var, model
     real(c_double) :: model_params(4)
     integer(c_int) :: symp6_tree
     integer(c_int) :: gw
     real(c_double) :: gw_mpl
     integer(c_int) :: reserved(4)
  end type lat_config

  integer, parameter, public :: LAT_F64 = 0, LAT_MIXED = 1
  integer, parameter, public :: LAT_LAP_SPECTRAL = 0, LAT_LAP_STENCIL = 1
  integer, parameter, public :: LAT_CONFORMAL = 0, LAT_COSMIC = 1
  integer, parameter, public :: LAT_MODEL_PHI4 = 0, LAT_MODEL_TWOFIELD = 1
  integer, parameter, public :: LAT_NE = 12

  type(c_ptr), save :: ctx = c_null_ptr
  integer, save :: gnx = 0, gny = 0, gnz = 0, gnf = 0, gtime = 0
  integer, save, public :: ns = 0
  ! energies of the last scalar_rho() (hamiltonian_cosmic.f90:335-338 kecur,
  ! pecur, gecur, rhocur)
  real(dl), save :: kecur = 0._dl, pecur = 0._dl, gecur = 0._dl, rhocur = 0._dl

  public :: lattice_init, lattice_finalize, lattice_put_state, lattice_get_state
  public :: laplacian_3d, init_spectrum_3d, spectrum_3d, crossspec_3d
  public :: Hamiltonian_Split, symp_o2step, symp2, symp4, symp6, symp8, calc_metric
  public :: get_scale_factor, get_hubble, kinetic_norm, grav_energy, lattice_energies
  public :: scalar_rho, kinetic_energy, grad_energy, potential_energy, grad_en, grad_en_wlap
  public :: discrete_make_output
  public :: spectrum, spectrum_bin, spectrum_nofft, spectrum_bin_nofft, crossspec_nofft
  public :: lattice_put_gw, lattice_get_gw, gw_spectrum

  interface
     integer(c_int) function c_lat_create(cfg, out) bind(C, name="lat_create")
       import :: c_int, c_ptr, lat_config
       type(lat_config), intent(in) :: cfg
       type(c_ptr), intent(out) :: out
     end function c_lat_create
     subroutine c_lat_destroy(h) bind(C, name="lat_destroy")
       import :: c_ptr
       type(c_ptr), value :: h
     end subroutine c_lat_destroy
     type(c_ptr) function c_lat_last_error(h) bind(C, name="lat_last_error")
       import :: c_ptr
       type(c_ptr), value :: h
     end function c_lat_last_error
     integer(c_int) function c_lat_laplacian_3d(h, n1, n2, n3, f, dk) bind(C, name="lat_laplacian_3d")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int), value :: n1, n2, n3
       real(c_double) :: f(*)
       real(c_double), value :: dk
     end function c_lat_laplacian_3d
     integer(c_int) function c_lat_num_bins(nx, ny, nz) bind(C, name="lat_num_bins")
       import :: c_int
       integer(c_int), value :: nx, ny, nz
     end function c_lat_num_bins
     integer(c_int) function c_lat_spectrum_3d(h, s, f, fk) bind(C, name="lat_spectrum_3d")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double) :: s(*), f(*)
       type(c_ptr), value :: fk
     end function c_lat_spectrum_3d
     integer(c_int) function c_lat_crossspec_3d(h, fk1, fk2, sre, sim) bind(C, name="lat_crossspec_3d")
       import :: c_int, c_ptr, c_double_complex, c_double
       type(c_ptr), value :: h
       complex(c_double_complex) :: fk1(*), fk2(*)
       real(c_double) :: sre(*), sim(*)
     end function c_lat_crossspec_3d
     integer(c_int) function c_lat_set_state(h, fld, fldp, ghost, yscl, ysclp) bind(C, name="lat_set_state")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double) :: fld(*), fldp(*)
       integer(c_int), value :: ghost
       real(c_double), value :: yscl, ysclp
     end function c_lat_set_state
     integer(c_int) function c_lat_get_state(h, fld, fldp, ghost, yscl, ysclp) bind(C, name="lat_get_state")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double) :: fld(*), fldp(*)
       integer(c_int), value :: ghost
       real(c_double) :: yscl, ysclp
     end function c_lat_get_state
     integer(c_int) function c_lat_get_scalars(h, yscl, ysclp) bind(C, name="lat_get_scalars")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double) :: yscl, ysclp
     end function c_lat_get_scalars
     integer(c_int) function c_lat_split(h, dt, hamind) bind(C, name="lat_split")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), value :: dt
       integer(c_int), value :: hamind
     end function c_lat_split
     integer(c_int) function c_lat_symp_o2step(h, dt, w1, w2) bind(C, name="lat_symp_o2step")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), value :: dt, w1, w2
     end function c_lat_symp_o2step
     integer(c_int) function c_lat_symp(h, order, dt, nsteps) bind(C, name="lat_symp")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int), value :: order, nsteps
       real(c_double), value :: dt
     end function c_lat_symp
     integer(c_int) function c_lat_calc_metric(h) bind(C, name="lat_calc_metric")
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function c_lat_calc_metric
     integer(c_int) function c_lat_energies(h, out) bind(C, name="lat_energies")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double) :: out(*)
     end function c_lat_energies
     integer(c_int) function c_lat_num_bins_discrete(n) bind(C, name="lat_num_bins_discrete")
       import :: c_int
       integer(c_int), value :: n
     end function c_lat_num_bins_discrete
     integer(c_int) function c_lat_spectrum_discrete(h, f, s, kind) bind(C, name="lat_spectrum_discrete")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double) :: f(*), s(*)
       integer(c_int), value :: kind
     end function c_lat_spectrum_discrete
     integer(c_int) function c_lat_spectrum_discrete_k(h, fk, s, kind) bind(C, name="lat_spectrum_discrete_k")
       import :: c_int, c_ptr, c_double, c_double_complex
       type(c_ptr), value :: h
       complex(c_double_complex) :: fk(*)
       real(c_double) :: s(*)
       integer(c_int), value :: kind
     end function c_lat_spectrum_discrete_k
     integer(c_int) function c_lat_crossspec_discrete_k(h, fk1, fk2, s1, s2) bind(C, name="lat_crossspec_discrete_k")
       import :: c_int, c_ptr, c_double, c_double_complex
       type(c_ptr), value :: h
       complex(c_double_complex) :: fk1(*), fk2(*)
       real(c_double) :: s1(*), s2(*)
     end function c_lat_crossspec_discrete_k
     integer(c_int) function c_lat_set_gw(h, hij, pij) bind(C, name="lat_set_gw")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double) :: hij(*), pij(*)
     end function c_lat_set_gw
     integer(c_int) function c_lat_get_gw(h, hij, pij) bind(C, name="lat_get_gw")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double) :: hij(*), pij(*)
     end function c_lat_get_gw
     integer(c_int) function c_lat_gw_spectrum(h, out) bind(C, name="lat_gw_spectrum")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double) :: out(*)
     end function c_lat_gw_spectrum
     integer(c_int) function c_lat_discrete_output(h, time, dt, mpl, seed, dir, vals, psd) &
          bind(C, name="lat_discrete_output")
       import :: c_int, c_ptr, c_double, c_char
       type(c_ptr), value :: h
       real(c_double), value :: time, dt, mpl
       integer(c_int), value :: seed
       character(kind=c_char) :: dir(*)
       type(c_ptr), value :: vals, psd
     end function c_lat_discrete_output
     integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
       import :: c_size_t, c_ptr
       type(c_ptr), value :: s
     end function c_strlen
  end interface

contains

  subroutine check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    type(c_ptr) :: p
    integer :: n
    if (rc == 0) return
    p = c_lat_last_error(ctx)
    n = int(c_strlen(p))
    call c_f_pointer(p, msg, [max(n, 1)])
    print *, where, " failed: status ", rc, " ", msg(1:n)
    stop
  end subroutine check

  ! Context creation: replaces the compile-time params.f90 / macros.h /
  ! HAM= model choice with runtime arguments.
  subroutine lattice_init(nx, ny, nz, nfld, len, precision, laplacian, time_var, model, model_params, stencil, &
                          gw, gw_mpl)
    integer, intent(in) :: nx, ny, nz, nfld, precision, laplacian, time_var, model
    real(dl), intent(in) :: len, model_params(:)
    real(dl), intent(in), optional :: stencil(5)
    logical, intent(in), optional :: gw        ! evolve hij, pi_hij (grav-waves/data-gw.f90)
    real(dl), intent(in), optional :: gw_mpl   ! M_p in program units (default 1)
    type(lat_config) :: cfg
    cfg%nx = nx; cfg%ny = ny; cfg%nz = nz; cfg%nfld = nfld
    cfg%len = len
    cfg%precision = precision; cfg%laplacian = laplacian
    cfg%stencil = (/ -6._dl, 1._dl, 0._dl, 0._dl, 1._dl /)
    if (present(stencil)) cfg%stencil = stencil
    cfg%time_var = time_var; cfg%model = model
    cfg%model_params = 0._dl
    cfg%model_params(1:size(model_params)) = model_params
    cfg%symp6_tree = 1
    cfg%gw = 0; cfg%gw_mpl = 0._dl
    if (present(gw)) then
       if (gw) cfg%gw = 1
    end if
    if (present(gw_mpl)) cfg%gw_mpl = gw_mpl
    cfg%reserved = 0
    if (c_associated(ctx)) call c_lat_destroy(ctx)
    call check(c_lat_create(cfg, ctx), "lat_create")
    gnx = nx; gny = ny; gnz = nz; gnf = nfld; gtime = time_var
  end subroutine lattice_init

  subroutine lattice_finalize()
    if (c_associated(ctx)) call c_lat_destroy(ctx)
    ctx = c_null_ptr
  end subroutine lattice_finalize

  ! fld(nfld, 0:nx+1, 0:ny+1, 0:nz+1) (ghost = 1) or fld(nfld, nx, ny, nz) (ghost = 0)
  subroutine lattice_put_state(fld, fldp, ghost, yscl, ysclp)
    real(dl), intent(in), target :: fld(*), fldp(*)
    integer, intent(in) :: ghost
    real(dl), intent(in) :: yscl, ysclp
    real(c_double), pointer :: a(:), b(:)
    call c_f_pointer(c_loc(fld), a, [1])
    call c_f_pointer(c_loc(fldp), b, [1])
    call check(c_lat_set_state(ctx, a, b, int(ghost, c_int), yscl, ysclp), "lat_set_state")
  end subroutine lattice_put_state

  subroutine lattice_get_state(fld, fldp, ghost, yscl, ysclp)
    real(dl), intent(out), target :: fld(*), fldp(*)
    integer, intent(in) :: ghost
    real(dl), intent(out) :: yscl, ysclp
    real(c_double), pointer :: a(:), b(:)
    call c_f_pointer(c_loc(fld), a, [1])
    call c_f_pointer(c_loc(fldp), b, [1])
    call check(c_lat_get_state(ctx, a, b, int(ghost, c_int), yscl, ysclp), "lat_get_state")
  end subroutine lattice_get_state

  ! fftw_mod.f90:177-207 (Fk, planf, planb ignored)
  subroutine laplacian_3d(n1, n2, n3, f, Fk, dk, planf, planb)
    integer :: n1, n2, n3
    real(c_double), target :: f(n1, n2, n3)
    complex(c_double_complex) :: Fk(*)
    real(dl) :: dk
    type(c_ptr) :: planf, planb
    call check(c_lat_laplacian_3d(ctx, int(n1, c_int), int(n2, c_int), int(n3, c_int), f, dk), "laplacian_3d")
  end subroutine laplacian_3d

  ! analysis_spec.f90:48-57
  subroutine init_spectrum_3d(nx, ny, nz)
    integer :: nx, ny, nz
    ns = int(c_lat_num_bins(int(nx, c_int), int(ny, c_int), int(nz, c_int)))
  end subroutine init_spectrum_3d

  ! analysis_spec.f90:59-102; Fk receives the R2C of f
  subroutine spectrum_3d(S, f, Fk, plan)
    real(dl) :: S(:)
    real(c_double) :: f(:,:,:)
    complex(c_double_complex), target :: Fk(:,:,:)
    type(c_ptr), optional :: plan
    call check(c_lat_spectrum_3d(ctx, S, f, c_loc(Fk)), "spectrum_3d")
  end subroutine spectrum_3d

  ! analysis_spec.f90:145-178
  subroutine crossspec_3d(Fk1, Fk2, S_real, S_imag)
    complex(c_double_complex) :: Fk1(:,:,:), Fk2(:,:,:)
    real(dl) :: S_real(:), S_imag(:)
    call check(c_lat_crossspec_3d(ctx, Fk1, Fk2, S_real, S_imag), "crossspec_3d")
  end subroutine crossspec_3d

  ! Discrete/Entropy/analysis.f90:89-114 (two-bin weights, /n^6); S(ns_d)
  subroutine spectrum(f, S)
    real(c_double) :: f(:,:,:)
    real(dl) :: S(:)
    call check(c_lat_spectrum_discrete(ctx, f, S, 0_c_int), "spectrum")
  end subroutine spectrum

  ! Discrete/Entropy/analysis.f90:22-52 (integer-count bins)
  subroutine spectrum_bin(f, S)
    real(c_double) :: f(:,:,:)
    real(dl) :: S(:)
    call check(c_lat_spectrum_discrete(ctx, f, S, 1_c_int), "spectrum_bin")
  end subroutine spectrum_bin

  ! Discrete/Entropy/analysis.f90:118-147
  subroutine spectrum_nofft(famp, S)
    complex(c_double_complex) :: famp(:,:,:)
    real(dl) :: S(:)
    call check(c_lat_spectrum_discrete_k(ctx, famp, S, 0_c_int), "spectrum_nofft")
  end subroutine spectrum_nofft

  ! Discrete/Entropy/analysis.f90:54-84 (the reference bins its module Fk here)
  subroutine spectrum_bin_nofft(famp, S)
    complex(c_double_complex) :: famp(:,:,:)
    real(dl) :: S(:)
    call check(c_lat_spectrum_discrete_k(ctx, famp, S, 1_c_int), "spectrum_bin_nofft")
  end subroutine spectrum_bin_nofft

  ! Discrete/Entropy/analysis.f90:149-186
  subroutine crossspec_nofft(famp1, famp2, S1, S2)
    complex(c_double_complex) :: famp1(:,:,:), famp2(:,:,:)
    real(dl) :: S1(:), S2(:)
    call check(c_lat_crossspec_discrete_k(ctx, famp1, famp2, S1, S2), "crossspec_nofft")
  end subroutine crossspec_nofft

  ! hij(2,3,n,n,n), pi_hij (grav-waves/data-gw.f90:14,28,33)
  subroutine lattice_put_gw(hij, pi_hij)
    real(dl) :: hij(:,:,:,:,:), pi_hij(:,:,:,:,:)
    call check(c_lat_set_gw(ctx, hij, pi_hij), "lattice_put_gw")
  end subroutine lattice_put_gw

  subroutine lattice_get_gw(hij, pi_hij)
    real(dl) :: hij(:,:,:,:,:), pi_hij(:,:,:,:,:)
    call check(c_lat_get_gw(ctx, hij, pi_hij), "lattice_get_gw")
  end subroutine lattice_get_gw

  ! TT-projected GW spectra: S_pi(ns) from pi_hij, S_h(ns) from hij
  subroutine gw_spectrum(S_pi, S_h)
    real(dl) :: S_pi(:), S_h(:)
    real(dl), allocatable :: out(:)
    integer :: nb
    nb = int(c_lat_num_bins(int(gnx, c_int), int(gny, c_int), int(gnz, c_int)))
    allocate(out(2 * nb))
    call check(c_lat_gw_spectrum(ctx, out), "gw_spectrum")
    S_pi(1:nb) = out(1:nb)
    S_h(1:nb) = out(nb + 1:2 * nb)
    deallocate(out)
  end subroutine gw_spectrum

  ! hamiltonian_conformal.f90:66-81 / hamiltonian_cosmic.f90:74-89
  subroutine Hamiltonian_Split(dt, hamind)
    real(dl) :: dt
    integer :: hamind
    call check(c_lat_split(ctx, dt, int(hamind, c_int)), "Hamiltonian_Split")
  end subroutine Hamiltonian_Split

  ! hamiltonian_conformal.f90:50-64
  subroutine symp_o2step(dt, w1, w2)
    real(dl) :: dt, w1, w2
    call check(c_lat_symp_o2step(ctx, dt, w1, w2), "symp_o2step")
  end subroutine symp_o2step

  ! mpi_dev/integrator.f90:11-110
  subroutine symp2(dt, nsteps)
    real(dl) :: dt
    integer :: nsteps
    call check(c_lat_symp(ctx, 2_c_int, dt, int(nsteps, c_int)), "symp2")
  end subroutine symp2

  subroutine symp4(dt, nsteps)
    real(dl) :: dt
    integer :: nsteps
    call check(c_lat_symp(ctx, 4_c_int, dt, int(nsteps, c_int)), "symp4")
  end subroutine symp4

  subroutine symp6(dt, nsteps)
    real(dl) :: dt
    integer :: nsteps
    call check(c_lat_symp(ctx, 6_c_int, dt, int(nsteps, c_int)), "symp6")
  end subroutine symp6

  subroutine symp8(dt, nsteps)
    real(dl) :: dt
    integer :: nsteps
    call check(c_lat_symp(ctx, 8_c_int, dt, int(nsteps, c_int)), "symp8")
  end subroutine symp8

  ! hamiltonian_conformal.f90:263-304 / hamiltonian_cosmic.f90:351-403
  subroutine calc_metric()
    call check(c_lat_calc_metric(ctx), "calc_metric")
  end subroutine calc_metric

  ! dump_rho / scalar_rho quantities: a, H, rho, KE, PE, GE, -3H^2, residual, ...
  subroutine lattice_energies(e)
    real(dl), intent(out) :: e(LAT_NE)
    call check(c_lat_energies(ctx, e), "energies")
  end subroutine lattice_energies

  ! hamiltonian_conformal.f90:242-261 / hamiltonian_cosmic.f90:234-253: scalar
  ! reads of the background (yscl, ysclp), no lattice work
  real(dl) function get_scale_factor()
    real(dl) :: y, p
    call check(c_lat_get_scalars(ctx, y, p), "get_scalars")
    if (gtime == LAT_CONFORMAL) then
       get_scale_factor = y
    else
       get_scale_factor = y**(2._dl / 3._dl)
    end if
  end function get_scale_factor

  real(dl) function get_hubble()
    real(dl) :: y, p
    call check(c_lat_get_scalars(ctx, y, p), "get_scalars")
    if (gtime == LAT_CONFORMAL) then
       get_hubble = -p / 6._dl / y**2
    else
       get_hubble = -p / y / 4._dl
    end if
  end function get_hubble

  real(dl) function grav_energy()
    grav_energy = -3._dl * get_hubble()**2
  end function grav_energy

  real(dl) function kinetic_norm()
    real(dl) :: y, p
    call check(c_lat_get_scalars(ctx, y, p), "get_scalars")
    ! 1/a^6 in both time variables: y^-6 (conformal, y = a), y^-4 (cosmic, y = a^{3/2})
    if (gtime == LAT_CONFORMAL) then
       kinetic_norm = 1._dl / y**6
    else
       kinetic_norm = 1._dl / y**4
    end if
  end function kinetic_norm

  ! mpi_dev/hamiltonian_cosmic.f90:268-349: rho = KE + PE + GE of the current
  ! state (one lattice reduction with the Laplacian); keeps the components
  ! for kinetic_energy / grad_energy / potential_energy (:255-266)
  real(dl) function scalar_rho()
    real(dl) :: e(LAT_NE)
    call lattice_energies(e)
    kecur = e(4); pecur = e(5); gecur = e(6); rhocur = e(3)
    scalar_rho = e(3)
  end function scalar_rho

  real(dl) function kinetic_energy()
    kinetic_energy = kecur
  end function kinetic_energy

  real(dl) function grad_energy()
    grad_energy = gecur
  end function grad_energy

  real(dl) function potential_energy()
    potential_energy = pecur
  end function potential_energy

  ! Discrete/Entropy/iofile.f90:317-335 make_output on a finite-difference
  ! context: appends the LOG.out line and the PSD block (head() headers on
  ! new files) in directory dir (default: the working directory); dt, mpl
  ! and seed go into the header's run line as in outputfile.f90:1231
  subroutine discrete_make_output(time, dt, mpl, seed, dir)
    real(dl), intent(in) :: time, dt, mpl
    integer, intent(in) :: seed
    character(*), intent(in), optional :: dir
    character(kind=c_char, len=:), allocatable :: d
    if (present(dir)) then
       d = trim(dir) // c_null_char
    else
       d = "." // c_null_char
    end if
    call check(c_lat_discrete_output(ctx, time, dt, mpl, int(seed, c_int), d, c_null_ptr, c_null_ptr), &
               "discrete_make_output")
  end subroutine discrete_make_output

  ! Discrete/Entropy/homlat.f90:88-115: gradient energy from STENCIL(c,GRAD2)
  ! (finite-difference contexts; NaN for the spectral Laplacian)
  real(dl) function grad_en()
    real(dl) :: e(LAT_NE)
    call lattice_energies(e)
    grad_en = e(12)
  end function grad_en

  ! homlat.f90:120-147: gradient energy from -phi STENCIL(c,LAPLACE) (the GE
  ! of dump_rho for the spectral Laplacian)
  real(dl) function grad_en_wlap()
    real(dl) :: e(LAT_NE)
    call lattice_energies(e)
    grad_en_wlap = e(6)
  end function grad_en_wlap

end module lattice_dropin
