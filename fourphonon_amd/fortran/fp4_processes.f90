! Drop-in replacements for the four-phonon procedures of FourPhonon's
! `module processes` (Src/processes.f90) and their three-phonon companions,
! running on an MI355X through the fp4 C ABI (include/fp4.h) via
! iso_c_binding.
!
! Procedure names, argument order, kinds and array shapes are those of the
! reference:
!   NP_driver_4ph                 Src/processes.f90:2749-2765
!   Ind_only_driver_4ph           Src/processes.f90:3080-3100
!   RTA_driver_4ph                Src/processes.f90:5174-5203
!   RTA_driver_4ph_GPU_using_Ind  Src/processes.f90:5315-5359
!   Ind_driver_4ph                Src/processes.f90:3551-3589
!   Vp_pp / Vp_pm / Vp_mm         Src/processes.f90:142-330
!   NP_driver                     Src/processes.f90:983-995
!   Ind_only_driver               Src/processes.f90:1198-1214
!   RTA_driver                    Src/processes.f90:1676-1697
!   RTA_driver_GPU_using_Ind      Src/processes.f90:1778-1808
!   Ind_driver                    Src/processes.f90:524-561
!   Vp_plus / Vp_minus            Src/processes.f90:25-139
!   iteration0                    Src/iterations.f90:13-30
!   iteration                     Src/iterations.f90:34-93
!   iteration_4ph                 Src/iterations.f90:95-120
!   TConduct                      Src/conductivity.f90:11-16
! The reference reads the crystal and run parameters from `module config`
! globals (Src/config.f90:9-51); here they are given once to fp4_init, which
! a caller invokes after read_config (and again after a change of T through
! fp4_set_temperature); the globals of the iterative solver (V, Nsymm_rot,
! symmetrizers, crotations and the ID_equi table of symmetry_map) are given
! to fp4_init_iterations.
!
! Deviation from the reference: an error in the device path is reported on
! error_unit and stops the program (error stop), where the reference prints
! and carries on.
module fp4_processes
  use iso_c_binding
  use iso_fortran_env, only: error_unit
  implicit none
  private

  integer(c_int32_t), parameter, public :: FP4_PREC_FP64 = 0, FP4_PREC_FP64_DENSE = 1, &
       FP4_PREC_FP32 = 2, FP4_PREC_BF16X3 = 3, FP4_PREC_FP16X3 = 4

  type, bind(C) :: fp4_config
     integer(c_int32_t) :: ngrid(3)
     integer(c_int32_t) :: nbands
     integer(c_int32_t) :: natoms
     integer(c_int32_t) :: nelements
     real(c_double) :: rlattvec(9)
     type(c_ptr) :: masses
     type(c_ptr) :: types
     real(c_double) :: T
     real(c_double) :: scalebroad
     real(c_double) :: omega_max
     integer(c_int32_t) :: onlyharmonic
     integer(c_int32_t) :: device
  end type fp4_config

  interface
     function c_last_error() bind(C, name="fp4_last_error")
       import :: c_ptr
       type(c_ptr) :: c_last_error
     end function c_last_error
     integer(c_int) function c_ctx_create(cfg, ctx) bind(C, name="fp4_ctx_create")
       import :: c_int, c_ptr, fp4_config
       type(fp4_config), intent(in) :: cfg
       type(c_ptr), intent(out) :: ctx
     end function c_ctx_create
     integer(c_int) function c_ctx_destroy(ctx) bind(C, name="fp4_ctx_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function c_ctx_destroy
     integer(c_int) function c_set_temperature(ctx, T) bind(C, name="fp4_ctx_set_temperature")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: T
     end function c_set_temperature
     integer(c_int) function c_set_modes(ctx, energy, velocity, eigenvect, IJK, nlist, List) &
          bind(C, name="fp4_ctx_set_modes")
       import :: c_int, c_ptr, c_double, c_double_complex, c_int32_t
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: energy(*), velocity(*)
       complex(c_double_complex), intent(in), optional :: eigenvect(*)   ! absent: counts and lists only
       integer(c_int32_t), intent(in) :: IJK(*), List(*)
       integer(c_int32_t), value :: nlist
     end function c_set_modes
     integer(c_int) function c_set_ifc4(ctx, ntri, Phi, R_j, R_k, R_l, Ii, Ij, Ik, Il) &
          bind(C, name="fp4_ctx_set_ifc4")
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: ntri
       real(c_double), intent(in) :: Phi(*), R_j(*), R_k(*), R_l(*)
       integer(c_int32_t), intent(in) :: Ii(*), Ij(*), Ik(*), Il(*)
     end function c_set_ifc4
     integer(c_int) function c_np(ctx, Npp, Ppp, Npm, Ppm, Nmm, Pmm) bind(C, name="fp4_np_4ph")
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), intent(out) :: Npp(*), Npm(*), Nmm(*)
       real(c_double), intent(out) :: Ppp(*), Ppm(*), Pmm(*)
     end function c_np
     integer(c_int) function c_np_sample(ctx, ns, seed, Npp, Ppp, Npm, Ppm, Nmm, Pmm) bind(C, name="fp4_np_4ph_sample")
       import :: c_int, c_ptr, c_double, c_int64_t, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: ns
       integer(c_int64_t), value :: seed
       integer(c_int64_t), intent(out) :: Npp(*), Npm(*), Nmm(*)
       real(c_double), intent(out) :: Ppp(*), Ppm(*), Pmm(*)
     end function c_np_sample
     integer(c_int) function c_rta_sample(ctx, ns, seed, r4, rpp, rpm, rmm, wpp, wpm, wmm) &
          bind(C, name="fp4_rta_4ph_sample")
       import :: c_int, c_ptr, c_double, c_int64_t, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: ns
       integer(c_int64_t), value :: seed
       real(c_double), intent(out) :: r4(*), rpp(*), rpm(*), rmm(*), wpp(*), wpm(*), wmm(*)
     end function c_rta_sample
     integer(c_int) function c_ind_only(ctx, Npp, Npm, Nmm, a2, a3, a4, b2, b3, b4, c2, c3, c4) &
          bind(C, name="fp4_ind_only_4ph")
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), intent(in) :: Npp(*), Npm(*), Nmm(*)
       integer(c_int64_t), intent(out) :: a2(*), a3(*), a4(*), b2(*), b3(*), b4(*), c2(*), c3(*), c4(*)
     end function c_ind_only
     integer(c_int) function c_autotune(ctx, prec, nm, modes, lo, hi, ib, tq, sl, ms) &
          bind(C, name="fp4_ctx_autotune_dense")
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: prec, nm, lo, hi
       integer(c_int32_t), intent(in) :: modes(*)
       integer(c_int32_t), intent(out) :: ib, tq, sl
       real(c_double), intent(out) :: ms
     end function c_autotune
     integer(c_int) function c_rta(ctx, prec, r4, rpp, rpm, rmm, wpp, wpm, wmm) bind(C, name="fp4_rta_4ph")
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: prec
       real(c_double), intent(out) :: r4(*), rpp(*), rpm(*), rmm(*), wpp(*), wpm(*), wmm(*)
     end function c_rta
     integer(c_int) function c_rta_ind(ctx, Npp, Npm, Nmm, a2, a3, a4, b2, b3, b4, c2, c3, c4, &
          r4, rpp, rpm, rmm, wpp, wpm, wmm) bind(C, name="fp4_rta_4ph_using_ind")
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), intent(in) :: Npp(*), Npm(*), Nmm(*)
       integer(c_int64_t), intent(in) :: a2(*), a3(*), a4(*), b2(*), b3(*), b4(*), c2(*), c3(*), c4(*)
       real(c_double), intent(out) :: r4(*), rpp(*), rpm(*), rmm(*), wpp(*), wpm(*), wmm(*)
     end function c_rta_ind
     integer(c_int) function c_ind(ctx, Npp, Npm, Nmm, a2, a3, a4, b2, b3, b4, c2, c3, c4, gpp, gpm, gmm, &
          r4, rpp, rpm, rmm, wpp, wpm, wmm) bind(C, name="fp4_ind_4ph")
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), intent(in) :: Npp(*), Npm(*), Nmm(*)
       integer(c_int64_t), intent(out) :: a2(*), a3(*), a4(*), b2(*), b3(*), b4(*), c2(*), c3(*), c4(*)
       real(c_double), intent(out) :: gpp(*), gpm(*), gmm(*)
       real(c_double), intent(out) :: r4(*), rpp(*), rpm(*), rmm(*), wpp(*), wpm(*), wmm(*)
     end function c_ind
     integer(c_int) function c_vp(ctx, channel, n, idx, out) bind(C, name="fp4_vp_4ph")
       import :: c_int, c_ptr, c_double, c_int32_t, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: channel
       integer(c_int64_t), value :: n
       integer(c_int32_t), intent(in) :: idx(*)
       real(c_double), intent(out) :: out(*)
     end function c_vp
     integer(c_int) function c_set_ifc3(ctx, ntri, Phi, R_j, R_k, Ii, Ij, Ik) bind(C, name="fp4_ctx_set_ifc3")
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: ntri
       real(c_double), intent(in) :: Phi(*), R_j(*), R_k(*)
       integer(c_int32_t), intent(in) :: Ii(*), Ij(*), Ik(*)
     end function c_set_ifc3
     integer(c_int) function c_np3(ctx, Np, Pp, Nm, Pm) bind(C, name="fp4_np_3ph")
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(out) :: Np(*), Nm(*)
       real(c_double), intent(out) :: Pp(*), Pm(*)
     end function c_np3
     integer(c_int) function c_np3_sample(ctx, ns, seed, Np, Pp, Nm, Pm) bind(C, name="fp4_np_3ph_sample")
       import :: c_int, c_ptr, c_double, c_int32_t, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: ns
       integer(c_int64_t), value :: seed
       integer(c_int32_t), intent(out) :: Np(*), Nm(*)
       real(c_double), intent(out) :: Pp(*), Pm(*)
     end function c_np3_sample
     integer(c_int) function c_rta3_sample(ctx, ns, seed, r, rp, rm, wp, wm) bind(C, name="fp4_rta_3ph_sample")
       import :: c_int, c_ptr, c_double, c_int32_t, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: ns
       integer(c_int64_t), value :: seed
       real(c_double), intent(out) :: r(*), rp(*), rm(*), wp(*), wm(*)
     end function c_rta3_sample
     integer(c_int) function c_ind_only3(ctx, Np, Nm, a2, a3, b2, b3) bind(C, name="fp4_ind_only_3ph")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(in) :: Np(*), Nm(*)
       integer(c_int32_t), intent(out) :: a2(*), a3(*), b2(*), b3(*)
     end function c_ind_only3
     integer(c_int) function c_rta3(ctx, r, rp, rm, wp, wm) bind(C, name="fp4_rta_3ph")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: r(*), rp(*), rm(*), wp(*), wm(*)
     end function c_rta3
     integer(c_int) function c_rta3_ind(ctx, Np, Nm, a2, b2, a3, b3, r, rp, rm, wp, wm) &
          bind(C, name="fp4_rta_3ph_using_ind")
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(in) :: Np(*), Nm(*), a2(*), b2(*), a3(*), b3(*)
       real(c_double), intent(out) :: r(*), rp(*), rm(*), wp(*), wm(*)
     end function c_rta3_ind
     integer(c_int) function c_ind3(ctx, Np, Nm, a2, a3, gp, b2, b3, gm, r, rp, rm, wp, wm) bind(C, name="fp4_ind_3ph")
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(in) :: Np(*), Nm(*)
       integer(c_int32_t), intent(out) :: a2(*), a3(*), b2(*), b3(*)
       real(c_double), intent(out) :: gp(*), gm(*), r(*), rp(*), rm(*), wp(*), wm(*)
     end function c_ind3
     integer(c_int) function c_vp3(ctx, channel, n, idx, out) bind(C, name="fp4_vp_3ph")
       import :: c_int, c_ptr, c_double, c_int32_t, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: channel
       integer(c_int64_t), value :: n
       integer(c_int32_t), intent(in) :: idx(*)
       real(c_double), intent(out) :: out(*)
     end function c_vp3
     integer(c_int) function c_set_symmetry(ctx, nsymm, ideq, symm, crot) bind(C, name="fp4_ctx_set_symmetry")
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: nsymm
       integer(c_int32_t), intent(in) :: ideq(*)
       real(c_double), intent(in) :: symm(*), crot(*)
     end function c_set_symmetry
     integer(c_int) function c_iteration0(ctx, nlist, Nequi, alleq, tau, F) bind(C, name="fp4_iteration0")
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: nlist
       integer(c_int32_t), intent(in) :: Nequi(*), alleq(*)
       real(c_double), intent(in) :: tau(*)
       real(c_double), intent(inout) :: F(*)
     end function c_iteration0
     integer(c_int) function c_set_operator(ctx, nlist, Nequi, alleq, types, Np, Nm, Ntp, Ntm, Npp, Npm, Nmm, &
          Ntpp, Ntpm, Ntmm, p2, p3, m2, m3, a2, a3, a4, b2, b3, b4, c2, c3, c4, gp, gm, gpp, gpm, gmm) &
          bind(C, name="fp4_bte_set_operator")
       import :: c_int, c_ptr, c_double, c_int32_t, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: nlist, Ntp, Ntm
       integer(c_int64_t), value :: Ntpp, Ntpm, Ntmm
       integer(c_int32_t), intent(in) :: Nequi(*), alleq(*), types(*), Np(*), Nm(*), p2(*), p3(*), m2(*), m3(*)
       integer(c_int64_t), intent(in) :: Npp(*), Npm(*), Nmm(*), a2(*), a3(*), a4(*), b2(*), b3(*), b4(*), &
            c2(*), c3(*), c4(*)
       real(c_double), intent(in) :: gp(*), gm(*), gpp(*), gpm(*), gmm(*)
     end function c_set_operator
     integer(c_int) function c_assemble(ctx, nlist, Nequi, alleq, types, prec, w3, w4, n3, n4, r3, r4) &
          bind(C, name="fp4_bte_assemble_operator")
       import :: c_int, c_ptr, c_double, c_int32_t, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: nlist, prec, w3, w4
       integer(c_int32_t), intent(in) :: Nequi(*), alleq(*), types(*)
       integer(c_int64_t), intent(out) :: n3, n4
       real(c_double), intent(out) :: r3(*), r4(*)
     end function c_assemble
     integer(c_int) function c_bte_iterate(ctx, tau, F) bind(C, name="fp4_bte_iterate")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: tau(*)
       real(c_double), intent(inout) :: F(*)
     end function c_bte_iterate
     integer(c_int) function c_tconduct(ctx, V, F, kappa, mode) bind(C, name="fp4_tconduct")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: V
       real(c_double), intent(in) :: F(*)
       real(c_double), intent(out) :: kappa(*), mode(*)
     end function c_tconduct
     integer(c_int) function c_bte_solve(ctx, V, tau, maxiter, eps, F, kappa, hist, rel, niter) &
          bind(C, name="fp4_bte_solve")
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: ctx
       real(c_double), value :: V, eps
       integer(c_int32_t), value :: maxiter
       real(c_double), intent(in) :: tau(*)
       real(c_double), intent(out) :: F(*), kappa(*), hist(*), rel(*)
       integer(c_int32_t), intent(out) :: niter
     end function c_bte_solve
  end interface

  type(c_ptr), save :: ctx = c_null_ptr
  integer(c_int32_t), save :: precision = FP4_PREC_FP64_DENSE
  ! num_sample_process_3ph[_phase_space] and num_sample_process_4ph[_phase_space]
  ! (Src/config.f90:23, default -1 = no sampling) and the seed of the
  ! sampling streams
  integer(kind=4), save :: num_sample_3ph_ = -1, num_sample_3ph_ps_ = -1
  integer(kind=4), save :: num_sample_4ph_ = -1, num_sample_4ph_ps_ = -1
  integer(c_int64_t), save :: sample_seed_ = 1
  integer(kind=4), save :: nptk_ = 0, nbands_ = 0
  real(c_double), save :: volume_ = 0.d0
  logical, save :: keep_operator_ = .false., operator_valid_ = .false.
  real(c_double), allocatable, target, save :: masses_(:)
  integer(c_int32_t), allocatable, target, save :: types_(:)

  public :: fp4_init, fp4_finalize, fp4_set_temperature, fp4_set_precision, fp4_set_sampling, fp4_autotune
  public :: NP_driver_4ph, Ind_only_driver_4ph, RTA_driver_4ph, RTA_driver_4ph_GPU_using_Ind, Ind_driver_4ph
  public :: Vp_pp, Vp_pm, Vp_mm
  public :: NP_driver, Ind_only_driver, RTA_driver, RTA_driver_GPU_using_Ind, Ind_driver, Vp_plus, Vp_minus
  public :: fp4_init_iterations, fp4_keep_operator, fp4_set_operator_4ph, fp4_solve_4ph, fp4_assemble_operator
  public :: iteration0, iteration, iteration_4ph, TConduct

contains

  subroutine check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: i
    if (rc == 0) return
    call c_f_pointer(c_last_error(), msg, [512])
    text = ""
    do i = 1, 512
       if (msg(i) == c_null_char) exit
       text(i:i) = msg(i)
    end do
    write(error_unit, "(a,a,a,i0,a,a)") "Error: ", where, " failed (code ", rc, "): ", trim(text)
    error stop 1
  end subroutine check

  ! The module globals of Src/config.f90:9-51 that the four-phonon path reads.
  subroutine fp4_init(ngrid, nbands, natoms, rlattvec, masses, types, T, scalebroad, omega_max, onlyharmonic, device)
    integer(kind=4), intent(in) :: ngrid(3), nbands, natoms
    real(kind=8), intent(in) :: rlattvec(3,3), masses(:), T, scalebroad, omega_max
    integer(kind=4), intent(in) :: types(:)
    logical, intent(in) :: onlyharmonic
    integer(kind=4), intent(in), optional :: device
    type(fp4_config) :: cfg
    if (c_associated(ctx)) call fp4_finalize()
    allocate(masses_(size(masses)), types_(size(types)))
    masses_ = masses
    types_ = types
    cfg%ngrid = ngrid
    cfg%nbands = nbands
    cfg%natoms = natoms
    cfg%nelements = size(masses)
    cfg%rlattvec = reshape(rlattvec, [9])
    cfg%masses = c_loc(masses_)
    cfg%types = c_loc(types_)
    cfg%T = T
    cfg%scalebroad = scalebroad
    cfg%omega_max = omega_max
    cfg%onlyharmonic = merge(1, 0, onlyharmonic)
    cfg%device = 0
    if (present(device)) cfg%device = device
    call check(c_ctx_create(cfg, ctx), "fp4_ctx_create")
    nptk_ = ngrid(1) * ngrid(2) * ngrid(3)
    nbands_ = nbands
  end subroutine fp4_init

  subroutine fp4_finalize()
    integer(c_int) :: rc
    if (c_associated(ctx)) rc = c_ctx_destroy(ctx)
    ctx = c_null_ptr
    if (allocated(masses_)) deallocate(masses_)
    if (allocated(types_)) deallocate(types_)
  end subroutine fp4_finalize

  subroutine fp4_set_temperature(T)
    real(kind=8), intent(in) :: T
    call check(c_set_temperature(ctx, T), "fp4_ctx_set_temperature")
  end subroutine fp4_set_temperature

  ! FP4_PREC_FP64 (per process, reference order), FP4_PREC_FP64_DENSE
  ! (default) or FP4_PREC_FP32 (dense, MFMA) for RTA_driver_4ph.
  subroutine fp4_set_precision(p)
    integer(kind=4), intent(in) :: p
    precision = p
  end subroutine fp4_set_precision

  ! The dense form's autotuner for the current precision, on every stride-th
  ! irreducible point (all its branches) over q' = 1..min(64, nptk); call it
  ! after the first driver call has uploaded the modes and constants.  The
  ! chosen shape is kept for the later RTA_driver_4ph calls.
  subroutine fp4_autotune(Nlist, stride)
    integer(kind=4), intent(in) :: Nlist, stride
    integer(c_int32_t), allocatable :: modes(:)
    integer(c_int32_t) :: ib, tq, sl
    real(c_double) :: ms
    integer :: ll, n
    allocate(modes((Nlist + max(stride, 1) - 1) / max(stride, 1)))
    n = 0
    do ll = 1, Nlist, max(stride, 1)
       n = n + 1
       modes(n) = (ll - 1) * nbands_ + 1
    end do
    call check(c_autotune(ctx, precision, n, modes, 1, min(64, nptk_), ib, tq, sl, ms), "fp4_autotune")
  end subroutine fp4_autotune

  ! The reference's four sampling switches, in the order of Src/config.f90:23
  ! (-1: the full drivers; n > 0: n sampled candidates per channel and mode),
  ! and the seed of the sampling streams (the reference reseeds from the
  ! clock; here a run is reproducible for a given seed).
  subroutine fp4_set_sampling(num_sample_process_3ph, num_sample_process_3ph_phase_space, &
       num_sample_process_4ph, num_sample_process_4ph_phase_space, seed)
    integer(kind=4), intent(in) :: num_sample_process_3ph, num_sample_process_3ph_phase_space
    integer(kind=4), intent(in) :: num_sample_process_4ph, num_sample_process_4ph_phase_space
    integer(kind=8), intent(in) :: seed
    num_sample_3ph_ = num_sample_process_3ph
    num_sample_3ph_ps_ = num_sample_process_3ph_phase_space
    num_sample_4ph_ = num_sample_process_4ph
    num_sample_4ph_ps_ = num_sample_process_4ph_phase_space
    sample_seed_ = seed
  end subroutine fp4_set_sampling

  subroutine upload_modes(energy, velocity, eigenvect, Nlist, List, IJK)
    real(kind=8), intent(in) :: energy(:,:), velocity(:,:,:)
    complex(kind=8), intent(in) :: eigenvect(:,:,:)
    integer(kind=4), intent(in) :: Nlist, List(:), IJK(:,:)
    call check(c_set_modes(ctx, energy, velocity, eigenvect, IJK, Nlist, List), "fp4_ctx_set_modes")
  end subroutine upload_modes

  subroutine upload_ifc4(Ntri, Phi, R_j, R_k, R_l, Index_i, Index_j, Index_k, Index_l)
    integer(kind=4), intent(in) :: Ntri
    real(kind=8), intent(in) :: Phi(:,:,:,:,:), R_j(:,:), R_k(:,:), R_l(:,:)
    integer(kind=4), intent(in) :: Index_i(:), Index_j(:), Index_k(:), Index_l(:)
    call check(c_set_ifc4(ctx, Ntri, Phi, R_j, R_k, R_l, Index_i, Index_j, Index_k, Index_l), "fp4_ctx_set_ifc4")
  end subroutine upload_ifc4

  ! NP_driver_4ph and Ind_only_driver_4ph have no eigenvectors: the mode data
  ! go up without them (a null pointer on the C side)
  subroutine upload_modes_noeig(energy, velocity, Nlist, List, IJK)
    real(kind=8), intent(in) :: energy(:,:), velocity(:,:,:)
    integer(kind=4), intent(in) :: Nlist, List(:), IJK(:,:)
    call check(c_set_modes(ctx, energy, velocity, IJK=IJK, nlist=Nlist, List=List), "fp4_ctx_set_modes")
  end subroutine upload_modes_noeig

  subroutine NP_driver_4ph(energy, velocity, Nlist, List, IJK, &
       N_plusplus, Pspace_plusplus_total, N_plusminus, Pspace_plusminus_total, &
       N_minusminus, Pspace_minusminus_total)
    real(kind=8), intent(in) :: energy(:,:), velocity(:,:,:)
    integer(kind=4), intent(in) :: Nlist, List(:), IJK(:,:)
    integer(kind=8), intent(out) :: N_plusplus(:), N_plusminus(:), N_minusminus(:)
    real(kind=8), intent(out) :: Pspace_plusplus_total(:,:), Pspace_plusminus_total(:,:), Pspace_minusminus_total(:,:)
    call upload_modes_noeig(energy, velocity, Nlist, List, IJK)
    if (num_sample_4ph_ps_ == -1) then
       call check(c_np(ctx, N_plusplus, Pspace_plusplus_total, N_plusminus, Pspace_plusminus_total, &
            N_minusminus, Pspace_minusminus_total), "NP_driver_4ph")
    else
       call check(c_np_sample(ctx, num_sample_4ph_ps_, sample_seed_, N_plusplus, Pspace_plusplus_total, &
            N_plusminus, Pspace_plusminus_total, N_minusminus, Pspace_minusminus_total), "NP_driver_4ph (sampled)")
    end if
  end subroutine NP_driver_4ph

  subroutine Ind_only_driver_4ph(energy, velocity, Nlist, List, IJK, N_plusplus, N_plusminus, N_minusminus, &
       Indof2ndPhonon_plusplus, Indof3rdPhonon_plusplus, Indof4thPhonon_plusplus, &
       Indof2ndPhonon_plusminus, Indof3rdPhonon_plusminus, Indof4thPhonon_plusminus, &
       Indof2ndPhonon_minusminus, Indof3rdPhonon_minusminus, Indof4thPhonon_minusminus)
    real(kind=8), intent(in) :: energy(:,:), velocity(:,:,:)
    integer(kind=4), intent(in) :: Nlist, List(:), IJK(:,:)
    integer(kind=8), intent(in) :: N_plusplus(:), N_plusminus(:), N_minusminus(:)
    integer(kind=8), intent(inout) :: Indof2ndPhonon_plusplus(:), Indof3rdPhonon_plusplus(:), Indof4thPhonon_plusplus(:)
    integer(kind=8), intent(inout) :: Indof2ndPhonon_plusminus(:), Indof3rdPhonon_plusminus(:), Indof4thPhonon_plusminus(:)
    integer(kind=8), intent(inout) :: Indof2ndPhonon_minusminus(:), Indof3rdPhonon_minusminus(:), &
         Indof4thPhonon_minusminus(:)
    call upload_modes_noeig(energy, velocity, Nlist, List, IJK)
    call check(c_ind_only(ctx, N_plusplus, N_plusminus, N_minusminus, &
         Indof2ndPhonon_plusplus, Indof3rdPhonon_plusplus, Indof4thPhonon_plusplus, &
         Indof2ndPhonon_plusminus, Indof3rdPhonon_plusminus, Indof4thPhonon_plusminus, &
         Indof2ndPhonon_minusminus, Indof3rdPhonon_minusminus, Indof4thPhonon_minusminus), "Ind_only_driver_4ph")
  end subroutine Ind_only_driver_4ph

  subroutine RTA_driver_4ph(energy, velocity, eigenvect, Nlist, List, IJK, &
       Ntri, Phi, R_j, R_k, R_l, Index_i, Index_j, Index_k, Index_l, &
       rate_scatt_4ph, rate_scatt_plusplus, rate_scatt_plusminus, rate_scatt_minusminus, &
       WP4_plusplus, WP4_plusminus, WP4_minusminus)
    real(kind=8), intent(in) :: energy(:,:), velocity(:,:,:)
    complex(kind=8), intent(in) :: eigenvect(:,:,:)
    integer(kind=4), intent(in) :: Nlist, List(:), IJK(:,:), Ntri
    real(kind=8), intent(in) :: Phi(:,:,:,:,:), R_j(:,:), R_k(:,:), R_l(:,:)
    integer(kind=4), intent(in) :: Index_i(:), Index_j(:), Index_k(:), Index_l(:)
    real(kind=8), intent(out) :: rate_scatt_4ph(:,:), rate_scatt_plusplus(:,:), rate_scatt_plusminus(:,:), &
         rate_scatt_minusminus(:,:), WP4_plusplus(:,:), WP4_plusminus(:,:), WP4_minusminus(:,:)
    call upload_modes(energy, velocity, eigenvect, Nlist, List, IJK)
    call upload_ifc4(Ntri, Phi, R_j, R_k, R_l, Index_i, Index_j, Index_k, Index_l)
    if (num_sample_4ph_ == -1) then
       call check(c_rta(ctx, precision, rate_scatt_4ph, rate_scatt_plusplus, rate_scatt_plusminus, &
            rate_scatt_minusminus, WP4_plusplus, WP4_plusminus, WP4_minusminus), "RTA_driver_4ph")
    else
       call check(c_rta_sample(ctx, num_sample_4ph_, sample_seed_, rate_scatt_4ph, rate_scatt_plusplus, &
            rate_scatt_plusminus, rate_scatt_minusminus, WP4_plusplus, WP4_plusminus, WP4_minusminus), &
            "RTA_driver_4ph (sampled)")
    end if
  end subroutine RTA_driver_4ph

  subroutine RTA_driver_4ph_GPU_using_Ind(energy, velocity, eigenvect, Nlist, List, IJK, &
       Ntri, Phi, R_j, R_k, R_l, Index_i, Index_j, Index_k, Index_l, &
       N_plusplus, N_plusminus, N_minusminus, &
       Indof2ndPhonon_plusplus, Indof3rdPhonon_plusplus, Indof4thPhonon_plusplus, &
       Indof2ndPhonon_plusminus, Indof3rdPhonon_plusminus, Indof4thPhonon_plusminus, &
       Indof2ndPhonon_minusminus, Indof3rdPhonon_minusminus, Indof4thPhonon_minusminus, &
       rate_scatt_4ph, rate_scatt_plusplus, rate_scatt_plusminus, rate_scatt_minusminus, &
       WP4_plusplus, WP4_plusminus, WP4_minusminus)
    real(kind=8), intent(in) :: energy(:,:), velocity(:,:,:)
    complex(kind=8), intent(in) :: eigenvect(:,:,:)
    integer(kind=4), intent(in) :: Nlist, List(:), IJK(:,:), Ntri
    real(kind=8), intent(in) :: Phi(:,:,:,:,:), R_j(:,:), R_k(:,:), R_l(:,:)
    integer(kind=4), intent(in) :: Index_i(:), Index_j(:), Index_k(:), Index_l(:)
    integer(kind=8), intent(in) :: N_plusplus(:), N_plusminus(:), N_minusminus(:)
    integer(kind=8), intent(in) :: Indof2ndPhonon_plusplus(:), Indof3rdPhonon_plusplus(:), Indof4thPhonon_plusplus(:)
    integer(kind=8), intent(in) :: Indof2ndPhonon_plusminus(:), Indof3rdPhonon_plusminus(:), Indof4thPhonon_plusminus(:)
    integer(kind=8), intent(in) :: Indof2ndPhonon_minusminus(:), Indof3rdPhonon_minusminus(:), &
         Indof4thPhonon_minusminus(:)
    real(kind=8), intent(out) :: rate_scatt_4ph(:,:), rate_scatt_plusplus(:,:), rate_scatt_plusminus(:,:), &
         rate_scatt_minusminus(:,:), WP4_plusplus(:,:), WP4_plusminus(:,:), WP4_minusminus(:,:)
    call upload_modes(energy, velocity, eigenvect, Nlist, List, IJK)
    call upload_ifc4(Ntri, Phi, R_j, R_k, R_l, Index_i, Index_j, Index_k, Index_l)
    call check(c_rta_ind(ctx, N_plusplus, N_plusminus, N_minusminus, &
         Indof2ndPhonon_plusplus, Indof3rdPhonon_plusplus, Indof4thPhonon_plusplus, &
         Indof2ndPhonon_plusminus, Indof3rdPhonon_plusminus, Indof4thPhonon_plusminus, &
         Indof2ndPhonon_minusminus, Indof3rdPhonon_minusminus, Indof4thPhonon_minusminus, &
         rate_scatt_4ph, rate_scatt_plusplus, rate_scatt_plusminus, rate_scatt_minusminus, &
         WP4_plusplus, WP4_plusminus, WP4_minusminus), "RTA_driver_4ph_GPU_using_Ind")
  end subroutine RTA_driver_4ph_GPU_using_Ind

  subroutine Ind_driver_4ph(energy, velocity, eigenvect, Nlist, List, IJK, &
       Ntri, Phi, R_j, R_k, R_l, Index_i, Index_j, Index_k, Index_l, &
       N_plusplus, N_plusminus, N_minusminus, &
       Indof2ndPhonon_plusplus, Indof3rdPhonon_plusplus, Indof4thPhonon_plusplus, &
       Indof2ndPhonon_plusminus, Indof3rdPhonon_plusminus, Indof4thPhonon_plusminus, &
       Indof2ndPhonon_minusminus, Indof3rdPhonon_minusminus, Indof4thPhonon_minusminus, &
       Gamma_plusplus, Gamma_plusminus, Gamma_minusminus, &
       rate_scatt_4ph, rate_scatt_plusplus, rate_scatt_plusminus, rate_scatt_minusminus, &
       WP4_plusplus, WP4_plusminus, WP4_minusminus)
    real(kind=8), intent(in) :: energy(:,:), velocity(:,:,:)
    complex(kind=8), intent(in) :: eigenvect(:,:,:)
    integer(kind=4), intent(in) :: Nlist, List(:), IJK(:,:), Ntri
    real(kind=8), intent(in) :: Phi(:,:,:,:,:), R_j(:,:), R_k(:,:), R_l(:,:)
    integer(kind=4), intent(in) :: Index_i(:), Index_j(:), Index_k(:), Index_l(:)
    integer(kind=8), intent(in) :: N_plusplus(:), N_plusminus(:), N_minusminus(:)
    integer(kind=8), intent(inout) :: Indof2ndPhonon_plusplus(:), Indof3rdPhonon_plusplus(:), Indof4thPhonon_plusplus(:)
    integer(kind=8), intent(inout) :: Indof2ndPhonon_plusminus(:), Indof3rdPhonon_plusminus(:), Indof4thPhonon_plusminus(:)
    integer(kind=8), intent(inout) :: Indof2ndPhonon_minusminus(:), Indof3rdPhonon_minusminus(:), &
         Indof4thPhonon_minusminus(:)
    real(kind=8), intent(inout) :: Gamma_plusplus(:), Gamma_plusminus(:), Gamma_minusminus(:)
    real(kind=8), intent(out) :: rate_scatt_4ph(:,:), rate_scatt_plusplus(:,:), rate_scatt_plusminus(:,:), &
         rate_scatt_minusminus(:,:), WP4_plusplus(:,:), WP4_plusminus(:,:), WP4_minusminus(:,:)
    call upload_modes(energy, velocity, eigenvect, Nlist, List, IJK)
    call upload_ifc4(Ntri, Phi, R_j, R_k, R_l, Index_i, Index_j, Index_k, Index_l)
    call check(c_ind(ctx, N_plusplus, N_plusminus, N_minusminus, &
         Indof2ndPhonon_plusplus, Indof3rdPhonon_plusplus, Indof4thPhonon_plusplus, &
         Indof2ndPhonon_plusminus, Indof3rdPhonon_plusminus, Indof4thPhonon_plusminus, &
         Indof2ndPhonon_minusminus, Indof3rdPhonon_minusminus, Indof4thPhonon_minusminus, &
         Gamma_plusplus, Gamma_plusminus, Gamma_minusminus, &
         rate_scatt_4ph, rate_scatt_plusplus, rate_scatt_plusminus, rate_scatt_minusminus, &
         WP4_plusplus, WP4_plusminus, WP4_minusminus), "Ind_driver_4ph")
  end subroutine Ind_driver_4ph

  ! One matrix element.  The modes and IFCs must have been uploaded by a
  ! driver call (or upload_* through a driver) on the same data; the q-vector
  ! arguments are accepted for interface compatibility, the device derives
  ! the exact phases from the grid indices.
  complex(kind=8) function vp_one(channel, i, j, k, l, q, qprime, qdprime, qtprime)
    integer(c_int32_t), intent(in) :: channel
    integer(kind=4), intent(in) :: i, j, k, l, q, qprime, qdprime, qtprime
    integer(c_int32_t) :: idx(8)
    real(c_double) :: out(2)
    idx = [i, j, k, l, q, qprime, qdprime, qtprime]
    call check(c_vp(ctx, channel, 1_c_int64_t, idx, out), "fp4_vp_4ph")
    vp_one = cmplx(out(1), out(2), kind=8)
  end function vp_one

  complex(kind=8) function Vp_pp(i, j, k, l, q, qprime, qdprime, qtprime, realqprime, realqdprime, realqtprime, &
       eigenvect, Ntri, Phi, R_j, R_k, R_l, Index_i, Index_j, Index_k, Index_l)
    integer(kind=4), intent(in) :: i, j, k, l, q, qprime, qdprime, qtprime, Ntri
    real(kind=8), intent(in) :: realqprime(3), realqdprime(3), realqtprime(3)
    complex(kind=8), intent(in) :: eigenvect(:,:,:)
    real(kind=8), intent(in) :: Phi(:,:,:,:,:), R_j(:,:), R_k(:,:), R_l(:,:)
    integer(kind=4), intent(in) :: Index_i(:), Index_j(:), Index_k(:), Index_l(:)
    Vp_pp = vp_one(0_c_int32_t, i, j, k, l, q, qprime, qdprime, qtprime)
  end function Vp_pp

  complex(kind=8) function Vp_pm(i, j, k, l, q, qprime, qdprime, qtprime, realqprime, realqdprime, realqtprime, &
       eigenvect, Ntri, Phi, R_j, R_k, R_l, Index_i, Index_j, Index_k, Index_l)
    integer(kind=4), intent(in) :: i, j, k, l, q, qprime, qdprime, qtprime, Ntri
    real(kind=8), intent(in) :: realqprime(3), realqdprime(3), realqtprime(3)
    complex(kind=8), intent(in) :: eigenvect(:,:,:)
    real(kind=8), intent(in) :: Phi(:,:,:,:,:), R_j(:,:), R_k(:,:), R_l(:,:)
    integer(kind=4), intent(in) :: Index_i(:), Index_j(:), Index_k(:), Index_l(:)
    Vp_pm = vp_one(1_c_int32_t, i, j, k, l, q, qprime, qdprime, qtprime)
  end function Vp_pm

  complex(kind=8) function Vp_mm(i, j, k, l, q, qprime, qdprime, qtprime, realqprime, realqdprime, realqtprime, &
       eigenvect, Ntri, Phi, R_j, R_k, R_l, Index_i, Index_j, Index_k, Index_l)
    integer(kind=4), intent(in) :: i, j, k, l, q, qprime, qdprime, qtprime, Ntri
    real(kind=8), intent(in) :: realqprime(3), realqdprime(3), realqtprime(3)
    complex(kind=8), intent(in) :: eigenvect(:,:,:)
    real(kind=8), intent(in) :: Phi(:,:,:,:,:), R_j(:,:), R_k(:,:), R_l(:,:)
    integer(kind=4), intent(in) :: Index_i(:), Index_j(:), Index_k(:), Index_l(:)
    Vp_mm = vp_one(2_c_int32_t, i, j, k, l, q, qprime, qdprime, qtprime)
  end function Vp_mm

  ! ------------------------------------------------------ three phonons
  subroutine upload_ifc3(Ntri, Phi, R_j, R_k, Index_i, Index_j, Index_k)
    integer(kind=4), intent(in) :: Ntri
    real(kind=8), intent(in) :: Phi(:,:,:,:), R_j(:,:), R_k(:,:)
    integer(kind=4), intent(in) :: Index_i(:), Index_j(:), Index_k(:)
    call check(c_set_ifc3(ctx, Ntri, Phi, R_j, R_k, Index_i, Index_j, Index_k), "fp4_ctx_set_ifc3")
  end subroutine upload_ifc3

  subroutine NP_driver(energy, velocity, Nlist, List, IJK, N_plus, Pspace_plus_total, N_minus, Pspace_minus_total)
    real(kind=8), intent(in) :: energy(:,:), velocity(:,:,:)
    integer(kind=4), intent(in) :: Nlist, List(:), IJK(:,:)
    integer(kind=4), intent(out) :: N_plus(:), N_minus(:)
    real(kind=8), intent(out) :: Pspace_plus_total(:,:), Pspace_minus_total(:,:)
    call upload_modes_noeig(energy, velocity, Nlist, List, IJK)
    if (num_sample_3ph_ps_ == -1) then
       call check(c_np3(ctx, N_plus, Pspace_plus_total, N_minus, Pspace_minus_total), "NP_driver")
    else
       call check(c_np3_sample(ctx, num_sample_3ph_ps_, sample_seed_, N_plus, Pspace_plus_total, N_minus, &
            Pspace_minus_total), "NP_driver (sampled)")
    end if
  end subroutine NP_driver

  subroutine Ind_only_driver(energy, velocity, Nlist, List, IJK, N_plus, N_minus, &
       Indof2ndPhonon_plus, Indof3rdPhonon_plus, Indof2ndPhonon_minus, Indof3rdPhonon_minus)
    real(kind=8), intent(in) :: energy(:,:), velocity(:,:,:)
    integer(kind=4), intent(in) :: Nlist, List(:), IJK(:,:)
    integer(kind=4), intent(in) :: N_plus(:), N_minus(:)
    integer(kind=4), intent(inout) :: Indof2ndPhonon_plus(:), Indof3rdPhonon_plus(:)
    integer(kind=4), intent(inout) :: Indof2ndPhonon_minus(:), Indof3rdPhonon_minus(:)
    call upload_modes_noeig(energy, velocity, Nlist, List, IJK)
    call check(c_ind_only3(ctx, N_plus, N_minus, Indof2ndPhonon_plus, Indof3rdPhonon_plus, &
         Indof2ndPhonon_minus, Indof3rdPhonon_minus), "Ind_only_driver")
  end subroutine Ind_only_driver

  subroutine RTA_driver(energy, velocity, eigenvect, Nlist, List, IJK, Ntri, Phi, R_j, R_k, &
       Index_i, Index_j, Index_k, rate_scatt, rate_scatt_plus, rate_scatt_minus, WP3_plus, WP3_minus)
    real(kind=8), intent(in) :: energy(:,:), velocity(:,:,:)
    complex(kind=8), intent(in) :: eigenvect(:,:,:)
    integer(kind=4), intent(in) :: Nlist, List(:), IJK(:,:), Ntri
    real(kind=8), intent(in) :: Phi(:,:,:,:), R_j(:,:), R_k(:,:)
    integer(kind=4), intent(in) :: Index_i(:), Index_j(:), Index_k(:)
    real(kind=8), intent(out) :: rate_scatt(:,:), rate_scatt_plus(:,:), rate_scatt_minus(:,:), &
         WP3_plus(:,:), WP3_minus(:,:)
    call upload_modes(energy, velocity, eigenvect, Nlist, List, IJK)
    call upload_ifc3(Ntri, Phi, R_j, R_k, Index_i, Index_j, Index_k)
    if (num_sample_3ph_ == -1) then
       call check(c_rta3(ctx, rate_scatt, rate_scatt_plus, rate_scatt_minus, WP3_plus, WP3_minus), "RTA_driver")
    else
       call check(c_rta3_sample(ctx, num_sample_3ph_, sample_seed_, rate_scatt, rate_scatt_plus, rate_scatt_minus, &
            WP3_plus, WP3_minus), "RTA_driver (sampled)")
    end if
  end subroutine RTA_driver

  subroutine RTA_driver_GPU_using_Ind(energy, velocity, eigenvect, Nlist, List, IJK, Ntri, Phi, R_j, R_k, &
       Index_i, Index_j, Index_k, N_plus, N_minus, Indof2ndPhonon_plus, Indof2ndPhonon_minus, &
       Indof3rdPhonon_plus, Indof3rdPhonon_minus, rate_scatt, rate_scatt_plus, rate_scatt_minus, &
       WP3_plus, WP3_minus)
    real(kind=8), intent(in) :: energy(:,:), velocity(:,:,:)
    complex(kind=8), intent(in) :: eigenvect(:,:,:)
    integer(kind=4), intent(in) :: Nlist, List(:), IJK(:,:), Ntri
    real(kind=8), intent(in) :: Phi(:,:,:,:), R_j(:,:), R_k(:,:)
    integer(kind=4), intent(in) :: Index_i(:), Index_j(:), Index_k(:)
    integer(kind=4), intent(in) :: N_plus(:), N_minus(:)
    integer(kind=4), intent(in) :: Indof2ndPhonon_plus(:), Indof2ndPhonon_minus(:)
    integer(kind=4), intent(in) :: Indof3rdPhonon_plus(:), Indof3rdPhonon_minus(:)
    real(kind=8), intent(out) :: rate_scatt(:,:), rate_scatt_plus(:,:), rate_scatt_minus(:,:), &
         WP3_plus(:,:), WP3_minus(:,:)
    call upload_modes(energy, velocity, eigenvect, Nlist, List, IJK)
    call upload_ifc3(Ntri, Phi, R_j, R_k, Index_i, Index_j, Index_k)
    call check(c_rta3_ind(ctx, N_plus, N_minus, Indof2ndPhonon_plus, Indof2ndPhonon_minus, &
         Indof3rdPhonon_plus, Indof3rdPhonon_minus, rate_scatt, rate_scatt_plus, rate_scatt_minus, &
         WP3_plus, WP3_minus), "RTA_driver_GPU_using_Ind")
  end subroutine RTA_driver_GPU_using_Ind

  subroutine Ind_driver(energy, velocity, eigenvect, Nlist, List, IJK, N_plus, N_minus, &
       Ntri, Phi, R_j, R_k, Index_i, Index_j, Index_k, &
       Indof2ndPhonon_plus, Indof3rdPhonon_plus, Gamma_plus, &
       Indof2ndPhonon_minus, Indof3rdPhonon_minus, Gamma_minus, &
       rate_scatt, rate_scatt_plus, rate_scatt_minus, WP3_plus, WP3_minus)
    real(kind=8), intent(in) :: energy(:,:), velocity(:,:,:)
    complex(kind=8), intent(in) :: eigenvect(:,:,:)
    integer(kind=4), intent(in) :: Nlist, List(:), IJK(:,:), Ntri
    integer(kind=4), intent(in) :: N_plus(:), N_minus(:)
    real(kind=8), intent(in) :: Phi(:,:,:,:), R_j(:,:), R_k(:,:)
    integer(kind=4), intent(in) :: Index_i(:), Index_j(:), Index_k(:)
    integer(kind=4), intent(out) :: Indof2ndPhonon_plus(:), Indof3rdPhonon_plus(:)
    integer(kind=4), intent(out) :: Indof2ndPhonon_minus(:), Indof3rdPhonon_minus(:)
    real(kind=8), intent(out) :: Gamma_plus(:), Gamma_minus(:)
    real(kind=8), intent(out) :: rate_scatt(:,:), rate_scatt_plus(:,:), rate_scatt_minus(:,:), &
         WP3_plus(:,:), WP3_minus(:,:)
    call upload_modes(energy, velocity, eigenvect, Nlist, List, IJK)
    call upload_ifc3(Ntri, Phi, R_j, R_k, Index_i, Index_j, Index_k)
    call check(c_ind3(ctx, N_plus, N_minus, Indof2ndPhonon_plus, Indof3rdPhonon_plus, Gamma_plus, &
         Indof2ndPhonon_minus, Indof3rdPhonon_minus, Gamma_minus, rate_scatt, rate_scatt_plus, &
         rate_scatt_minus, WP3_plus, WP3_minus), "Ind_driver")
  end subroutine Ind_driver

  complex(kind=8) function vp3_one(channel, i, j, k, q, qprime, qdprime)
    integer(c_int32_t), intent(in) :: channel
    integer(kind=4), intent(in) :: i, j, k, q, qprime, qdprime
    integer(c_int32_t) :: idx(6)
    real(c_double) :: out(2)
    idx = [i, j, k, q, qprime, qdprime]
    call check(c_vp3(ctx, channel, 1_c_int64_t, idx, out), "fp4_vp_3ph")
    vp3_one = cmplx(out(1), out(2), kind=8)
  end function vp3_one

  ! Vp_plus / Vp_minus use the modes and third-order clusters of the last
  ! three-phonon driver call; their array arguments are accepted for
  ! interface compatibility.
  complex(kind=8) function Vp_plus(i, j, k, q, qprime, qdprime, realqprime, realqdprime, eigenvect, &
       Ntri, Phi, R_j, R_k, Index_i, Index_j, Index_k)
    integer(kind=4), intent(in) :: i, j, k, q, qprime, qdprime, Ntri
    real(kind=8), intent(in) :: realqprime(3), realqdprime(3)
    complex(kind=8), intent(in) :: eigenvect(:,:,:)
    real(kind=8), intent(in) :: Phi(:,:,:,:), R_j(:,:), R_k(:,:)
    integer(kind=4), intent(in) :: Index_i(:), Index_j(:), Index_k(:)
    Vp_plus = vp3_one(0_c_int32_t, i, j, k, q, qprime, qdprime)
  end function Vp_plus

  complex(kind=8) function Vp_minus(i, j, k, q, qprime, qdprime, realqprime, realqdprime, eigenvect, &
       Ntri, Phi, R_j, R_k, Index_i, Index_j, Index_k)
    integer(kind=4), intent(in) :: i, j, k, q, qprime, qdprime, Ntri
    real(kind=8), intent(in) :: realqprime(3), realqdprime(3)
    complex(kind=8), intent(in) :: eigenvect(:,:,:)
    real(kind=8), intent(in) :: Phi(:,:,:,:), R_j(:,:), R_k(:,:)
    integer(kind=4), intent(in) :: Index_i(:), Index_j(:), Index_k(:)
    Vp_minus = vp3_one(1_c_int32_t, i, j, k, q, qprime, qdprime)
  end function Vp_minus

  ! ---------------------------------------------------------- iterations
  ! Globals read by Src/iterations.f90 and Src/conductivity.f90: the cell
  ! volume V (nm^3), Nsymm_rot, ID_equi from symmetry_map
  ! (Src/config.f90:409-431), symmetrizers(3,3,nptk) and crotations.
  subroutine fp4_init_iterations(V, Nsymm_rot, ID_equi, symmetrizers, crotations)
    real(kind=8), intent(in) :: V, symmetrizers(:,:,:), crotations(:,:,:)
    integer(kind=4), intent(in) :: Nsymm_rot, ID_equi(:,:)
    volume_ = V
    operator_valid_ = .false.
    call check(c_set_symmetry(ctx, Nsymm_rot, ID_equi, symmetrizers, crotations), "fp4_ctx_set_symmetry")
  end subroutine fp4_init_iterations

  ! With flag .true., the scattering operator (lists and Gamma) uploaded by
  ! the next iteration_4ph call stays on the device and later calls reuse it
  ! without reading their list and Gamma arguments again; the caller
  ! promises not to change them until fp4_keep_operator(.false.).  Default
  ! .false.: every call uploads its arguments, as the reference reads them.
  subroutine fp4_keep_operator(flag)
    logical, intent(in) :: flag
    keep_operator_ = flag
    operator_valid_ = .false.
  end subroutine fp4_keep_operator

  ! iteration0 and iteration_4ph use the omega/velocity resident on the
  ! device, i.e. those of the last driver call on this context; their omega
  ! and velocity arguments are accepted for interface compatibility.
  subroutine iteration0(Nlist, Nequi, ALLEquiList, omega, velocity, tau_zero, F_n)
    integer(kind=4), intent(in) :: Nlist, Nequi(:), ALLEquiList(:,:)
    real(kind=8), intent(in) :: omega(:,:), velocity(:,:,:), tau_zero(:,:)
    real(kind=8), intent(out) :: F_n(:,:,:)
    call check(c_iteration0(ctx, Nlist, Nequi, ALLEquiList, tau_zero, F_n), "iteration0")
  end subroutine iteration0

  subroutine fp4_set_operator_4ph(Nlist, Nequi, ALLEquiList, TypeofSymmetry, N_plus, N_minus, Ntotal_plus, &
       Ntotal_minus, N_plusplus, N_plusminus, N_minusminus, Ntotal_plusplus, Ntotal_plusminus, Ntotal_minusminus, &
       Indof2ndPhonon_plus, Indof3rdPhonon_plus, Indof2ndPhonon_minus, Indof3rdPhonon_minus, &
       Indof2ndPhonon_plusplus, Indof3rdPhonon_plusplus, Indof4thPhonon_plusplus, &
       Indof2ndPhonon_plusminus, Indof3rdPhonon_plusminus, Indof4thPhonon_plusminus, &
       Indof2ndPhonon_minusminus, Indof3rdPhonon_minusminus, Indof4thPhonon_minusminus, &
       Gamma_plus, Gamma_minus, Gamma_plusplus, Gamma_plusminus, Gamma_minusminus)
    integer(kind=4), intent(in) :: Nlist, Nequi(:), ALLEquiList(:,:), TypeofSymmetry(:,:)
    integer(kind=4), intent(in) :: N_plus(:), N_minus(:), Ntotal_plus, Ntotal_minus
    integer(kind=4), intent(in) :: Indof2ndPhonon_plus(:), Indof3rdPhonon_plus(:)
    integer(kind=4), intent(in) :: Indof2ndPhonon_minus(:), Indof3rdPhonon_minus(:)
    integer(kind=8), intent(in) :: N_plusplus(:), N_plusminus(:), N_minusminus(:)
    integer(kind=8), intent(in) :: Ntotal_plusplus, Ntotal_plusminus, Ntotal_minusminus
    integer(kind=8), intent(in) :: Indof2ndPhonon_plusplus(:), Indof3rdPhonon_plusplus(:), Indof4thPhonon_plusplus(:)
    integer(kind=8), intent(in) :: Indof2ndPhonon_plusminus(:), Indof3rdPhonon_plusminus(:), Indof4thPhonon_plusminus(:)
    integer(kind=8), intent(in) :: Indof2ndPhonon_minusminus(:), Indof3rdPhonon_minusminus(:), &
         Indof4thPhonon_minusminus(:)
    real(kind=8), intent(in) :: Gamma_plus(:), Gamma_minus(:), Gamma_plusplus(:), Gamma_plusminus(:), &
         Gamma_minusminus(:)
    call check(c_set_operator(ctx, Nlist, Nequi, ALLEquiList, TypeofSymmetry, N_plus, N_minus, Ntotal_plus, &
         Ntotal_minus, N_plusplus, N_plusminus, N_minusminus, Ntotal_plusplus, Ntotal_plusminus, Ntotal_minusminus, &
         Indof2ndPhonon_plus, Indof3rdPhonon_plus, Indof2ndPhonon_minus, Indof3rdPhonon_minus, &
         Indof2ndPhonon_plusplus, Indof3rdPhonon_plusplus, Indof4thPhonon_plusplus, &
         Indof2ndPhonon_plusminus, Indof3rdPhonon_plusminus, Indof4thPhonon_plusminus, &
         Indof2ndPhonon_minusminus, Indof3rdPhonon_minusminus, Indof4thPhonon_minusminus, &
         Gamma_plus, Gamma_minus, Gamma_plusplus, Gamma_plusminus, Gamma_minusminus), "fp4_bte_set_operator")
    operator_valid_ = .true.
  end subroutine fp4_set_operator_4ph

  subroutine iteration_4ph(Nlist, Nequi, ALLEquiList, TypeofSymmetry, N_plus, N_minus, Ntotal_plus, Ntotal_minus, &
       N_plusplus, N_plusminus, N_minusminus, Ntotal_plusplus, Ntotal_plusminus, Ntotal_minusminus, &
       Indof2ndPhonon_plus, Indof3rdPhonon_plus, Indof2ndPhonon_minus, Indof3rdPhonon_minus, &
       Indof2ndPhonon_plusplus, Indof3rdPhonon_plusplus, Indof4thPhonon_plusplus, &
       Indof2ndPhonon_plusminus, Indof3rdPhonon_plusminus, Indof4thPhonon_plusminus, &
       Indof2ndPhonon_minusminus, Indof3rdPhonon_minusminus, Indof4thPhonon_minusminus, &
       omega, velocity, Gamma_plus, Gamma_minus, Gamma_plusplus, Gamma_plusminus, Gamma_minusminus, tau_zero, F_n)
    integer(kind=4), intent(in) :: Nlist, Nequi(:), ALLEquiList(:,:), TypeofSymmetry(:,:)
    integer(kind=4), intent(in) :: N_plus(:), N_minus(:), Ntotal_plus, Ntotal_minus
    integer(kind=4), intent(in) :: Indof2ndPhonon_plus(:), Indof3rdPhonon_plus(:)
    integer(kind=4), intent(in) :: Indof2ndPhonon_minus(:), Indof3rdPhonon_minus(:)
    integer(kind=8), intent(in) :: N_plusplus(:), N_plusminus(:), N_minusminus(:)
    integer(kind=8), intent(in) :: Ntotal_plusplus, Ntotal_plusminus, Ntotal_minusminus
    integer(kind=8), intent(in) :: Indof2ndPhonon_plusplus(:), Indof3rdPhonon_plusplus(:), Indof4thPhonon_plusplus(:)
    integer(kind=8), intent(in) :: Indof2ndPhonon_plusminus(:), Indof3rdPhonon_plusminus(:), Indof4thPhonon_plusminus(:)
    integer(kind=8), intent(in) :: Indof2ndPhonon_minusminus(:), Indof3rdPhonon_minusminus(:), &
         Indof4thPhonon_minusminus(:)
    real(kind=8), intent(in) :: omega(:,:), velocity(:,:,:)
    real(kind=8), intent(in) :: Gamma_plus(:), Gamma_minus(:), Gamma_plusplus(:), Gamma_plusminus(:), &
         Gamma_minusminus(:)
    real(kind=8), intent(in) :: tau_zero(:,:)
    real(kind=8), intent(inout) :: F_n(:,:,:)
    if (.not. (keep_operator_ .and. operator_valid_)) then
       call fp4_set_operator_4ph(Nlist, Nequi, ALLEquiList, TypeofSymmetry, N_plus, N_minus, Ntotal_plus, &
            Ntotal_minus, N_plusplus, N_plusminus, N_minusminus, Ntotal_plusplus, Ntotal_plusminus, &
            Ntotal_minusminus, Indof2ndPhonon_plus, Indof3rdPhonon_plus, Indof2ndPhonon_minus, &
            Indof3rdPhonon_minus, Indof2ndPhonon_plusplus, Indof3rdPhonon_plusplus, Indof4thPhonon_plusplus, &
            Indof2ndPhonon_plusminus, Indof3rdPhonon_plusminus, Indof4thPhonon_plusminus, &
            Indof2ndPhonon_minusminus, Indof3rdPhonon_minusminus, Indof4thPhonon_minusminus, &
            Gamma_plus, Gamma_minus, Gamma_plusplus, Gamma_plusminus, Gamma_minusminus)
    end if
    call check(c_bte_iterate(ctx, tau_zero, F_n), "iteration_4ph")
  end subroutine iteration_4ph

  ! The three-phonon iteration of the reference's default mode
  ! (four_phonon_iteration = .false., Src/iterations.f90:34-93): the same
  ! in-place update with empty four-phonon lists; 4ph scattering enters only
  ! through tau_zero.
  subroutine iteration(Nlist, Nequi, ALLEquiList, TypeofSymmetry, N_plus, N_minus, Ntotal_plus, Ntotal_minus, &
       Indof2ndPhonon_plus, Indof3rdPhonon_plus, Indof2ndPhonon_minus, Indof3rdPhonon_minus, omega, &
       velocity, Gamma_plus, Gamma_minus, tau_zero, F_n)
    integer(kind=4), intent(in) :: Nlist, Nequi(:), ALLEquiList(:,:), TypeofSymmetry(:,:)
    integer(kind=4), intent(in) :: N_plus(:), N_minus(:), Ntotal_plus, Ntotal_minus
    integer(kind=4), intent(in) :: Indof2ndPhonon_plus(:), Indof3rdPhonon_plus(:)
    integer(kind=4), intent(in) :: Indof2ndPhonon_minus(:), Indof3rdPhonon_minus(:)
    real(kind=8), intent(in) :: omega(:,:), velocity(:,:,:)
    real(kind=8), intent(in) :: Gamma_plus(:), Gamma_minus(:), tau_zero(:,:)
    real(kind=8), intent(inout) :: F_n(:,:,:)
    integer(kind=8), allocatable :: zn(:), zl(:)
    real(kind=8), allocatable :: zg(:)
    if (.not. (keep_operator_ .and. operator_valid_)) then
       allocate(zn(size(N_plus)), zl(1), zg(1))
       zn = 0
       zl = 1
       zg = 0.d0
       call fp4_set_operator_4ph(Nlist, Nequi, ALLEquiList, TypeofSymmetry, N_plus, N_minus, Ntotal_plus, &
            Ntotal_minus, zn, zn, zn, 0_8, 0_8, 0_8, Indof2ndPhonon_plus, Indof3rdPhonon_plus, &
            Indof2ndPhonon_minus, Indof3rdPhonon_minus, zl, zl, zl, zl, zl, zl, zl, zl, zl, &
            Gamma_plus, Gamma_minus, zg, zg, zg)
    end if
    call check(c_bte_iterate(ctx, tau_zero, F_n), "iteration")
  end subroutine iteration

  ! TConduct at the context's T and the V of fp4_init_iterations.
  subroutine TConduct(omega, velocity, F_n, ThConductivity, ThConductivityMode)
    real(kind=8), intent(in) :: omega(:,:), velocity(:,:,:), F_n(:,:,:)
    real(kind=8), intent(out) :: ThConductivity(:,:,:), ThConductivityMode(:,:,:,:)
    call check(c_tconduct(ctx, volume_, F_n, ThConductivity, ThConductivityMode), "TConduct")
  end subroutine TConduct

  ! The 3ph+4ph operator of iteration_4ph assembled on the device at the
  ! current temperature, aggregated by partner mode instead of stored as
  ! Ind_driver / Ind_driver_4ph lists (which do not fit memory at large
  ! meshes), with the 4ph Gamma in the precision of fp4_set_precision and
  ! the 3ph Gamma in FP64.  rate_3ph and rate_4ph (Nb,Nlist) receive the
  ! rates of the same processes (for tau_zero).  Later iteration_4ph calls
  ! and fp4_solve_4ph use this operator (as with fp4_keep_operator(.true.))
  ! until the next fp4_set_operator_4ph or fp4_keep_operator call.
  subroutine fp4_assemble_operator(Nlist, Nequi, ALLEquiList, TypeofSymmetry, rate_3ph, rate_4ph)
    integer(kind=4), intent(in) :: Nlist, Nequi(:), ALLEquiList(:,:), TypeofSymmetry(:,:)
    real(kind=8), intent(out) :: rate_3ph(:,:), rate_4ph(:,:)
    integer(c_int64_t) :: n3, n4
    call check(c_assemble(ctx, Nlist, Nequi, ALLEquiList, TypeofSymmetry, precision, 1_c_int32_t, 1_c_int32_t, &
         n3, n4, rate_3ph, rate_4ph), "fp4_assemble_operator")
    keep_operator_ = .true.
    operator_valid_ = .true.
  end subroutine fp4_assemble_operator

  ! The convergence loop of Src/ShengBTE.f90:1250-1307 on the device, with
  ! the operator of the last fp4_set_operator_4ph / iteration_4ph call:
  ! kappa_iter(3,3,0:maxiter) and relchange(0:maxiter) receive entries
  ! 0..niter.
  subroutine fp4_solve_4ph(tau_zero, maxiter, eps, F_n, ThConductivity, kappa_iter, relchange, niter)
    real(kind=8), intent(in) :: tau_zero(:,:), eps
    integer(kind=4), intent(in) :: maxiter
    real(kind=8), intent(out) :: F_n(:,:,:), ThConductivity(:,:,:), kappa_iter(:,:,0:), relchange(0:)
    integer(kind=4), intent(out) :: niter
    if (size(kappa_iter, 3) < maxiter + 1 .or. size(relchange) < maxiter + 1) then
       write(error_unit, "(a)") "Error: fp4_solve_4ph needs kappa_iter(3,3,0:maxiter) and relchange(0:maxiter)"
       error stop 1
    end if
    call check(c_bte_solve(ctx, volume_, tau_zero, maxiter, eps, F_n, ThConductivity, kappa_iter, relchange, &
         niter), "fp4_bte_solve")
  end subroutine fp4_solve_4ph

end module fp4_processes
