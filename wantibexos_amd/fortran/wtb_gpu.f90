! ISO_C_BINDING interface to libwtb_bse.so (include/wtb_bse.h).
!
! The derived types mirror wtb_tb_model / wtb_bse_params field by field; the
! interfaces bind the C entry points one to one.  The drop-in subroutines with
! the reference's names and argument lists are in bse_solver_gpu.f90.
module wtb_gpu
	use iso_c_binding
	implicit none

	integer(c_int), parameter :: WTB_OK = 0
	integer(c_int), parameter :: WTB_FP64 = 0, WTB_MIXED = 1
	integer(c_int), parameter :: WTB_TA_FA = 0, WTB_TA_VE = 1, WTB_TA_BE = 2

	type, bind(C) :: wtb_tb_model
		integer(c_int) :: nw, nvec
		type(c_ptr) :: rlat, rvec, ffactor, hop, ihop
		real(c_double) :: efermi, scs
	end type

	type, bind(C) :: wtb_bse_params
		integer(c_int) :: ngrid(3)
		real(c_double) :: mshift(3)
		integer(c_int) :: nc, nv, coultype
		real(c_double) :: ediel(3)
		real(c_double) :: lc, ez, w, r0, ktol, cshift, ebse0, ebsef
		integer(c_int) :: numbse, dtfull, cpol, renorm, f154, precision
		integer(c_int) :: bset, ta
		real(c_double) :: temp, st, phavg
	end type

	interface
		integer(c_int) function wtb_create(device, ctx) bind(C, name="wtb_create")
			import :: c_int, c_ptr
			integer(c_int), value :: device
			type(c_ptr) :: ctx
		end function
		integer(c_int) function wtb_destroy(ctx) bind(C, name="wtb_destroy")
			import :: c_int, c_ptr
			type(c_ptr), value :: ctx
		end function
		type(c_ptr) function wtb_last_error() bind(C, name="wtb_last_error")
			import :: c_ptr
		end function
		integer(c_int) function wtb_eigsys(ctx, model, nk, kpt, energias, autovetores, nocp) &
				bind(C, name="wtb_eigsys")
			import :: c_int, c_ptr, c_double, wtb_tb_model
			type(c_ptr), value :: ctx
			type(wtb_tb_model) :: model
			integer(c_int), value :: nk
			real(c_double) :: kpt(*), energias(*)
			complex(c_double_complex) :: autovetores(*)
			integer(c_int) :: nocp(*)
		end function
		integer(c_int) function wtb_bsesolver(ctx, model, prm, w, act, eps, egap) &
				bind(C, name="wtb_bsesolver")
			import :: c_int, c_ptr, c_double, wtb_tb_model, wtb_bse_params
			type(c_ptr), value :: ctx
			type(wtb_tb_model) :: model
			type(wtb_bse_params) :: prm
			real(c_double) :: w(*), act(*)
			type(c_ptr), value :: eps
			real(c_double) :: egap
		end function
		integer(c_int) function wtb_bsesolver_wf(ctx, model, prm, w, act, eps, egap, wf_first, &
				wf_count, wf, kpt, nocp) bind(C, name="wtb_bsesolver_wf")
			import :: c_int, c_ptr, c_double, c_double_complex, wtb_tb_model, wtb_bse_params
			type(c_ptr), value :: ctx
			type(wtb_tb_model) :: model
			type(wtb_bse_params) :: prm
			real(c_double) :: w(*), act(*)
			type(c_ptr), value :: eps
			real(c_double) :: egap
			integer(c_int), value :: wf_first, wf_count
			complex(c_double_complex) :: wf(*)
			real(c_double) :: kpt(*)
			integer(c_int) :: nocp(*)
		end function
		integer(c_int) function wtb_bsebnds(ctx, model, prm, nq, qauxv, lanes, exk) &
				bind(C, name="wtb_bsebnds")
			import :: c_int, c_ptr, c_double, wtb_tb_model, wtb_bse_params
			type(c_ptr), value :: ctx
			type(wtb_tb_model) :: model
			type(wtb_bse_params) :: prm
			integer(c_int), value :: nq, lanes
			real(c_double) :: qauxv(*), exk(*)
		end function
		integer(c_int) function wtb_bsebnds_wf(ctx, model, prm, nq, qauxv, lanes, exk, wf_first, &
				wf_count, wf, kpt, nocp) bind(C, name="wtb_bsebnds_wf")
			import :: c_int, c_ptr, c_double, c_double_complex, wtb_tb_model, wtb_bse_params
			type(c_ptr), value :: ctx
			type(wtb_tb_model) :: model
			type(wtb_bse_params) :: prm
			integer(c_int), value :: nq, lanes, wf_first, wf_count
			real(c_double) :: qauxv(*), exk(*)
			complex(c_double_complex) :: wf(*)
			real(c_double) :: kpt(*)
			integer(c_int) :: nocp(*)
		end function
		integer(c_int) function wtb_sp_table(ctx, model, prm, vecres) bind(C, name="wtb_sp_table")
			import :: c_int, c_ptr, c_float, wtb_tb_model, wtb_bse_params
			type(c_ptr), value :: ctx
			type(wtb_tb_model) :: model
			type(wtb_bse_params) :: prm
			real(c_float) :: vecres(*)
		end function
		integer(c_int) function wtb_bse_spectrum(ctx, n, e, act6, rlat, ngrid, ebse0, ebsef, &
				numbse, sme, renorm, eps) bind(C, name="wtb_bse_spectrum")
			import :: c_int, c_ptr, c_double
			type(c_ptr), value :: ctx
			integer(c_int), value :: n
			real(c_double) :: e(*), act6(*), rlat(*)
			integer(c_int) :: ngrid(3)
			real(c_double), value :: ebse0, ebsef
			integer(c_int), value :: numbse
			real(c_double), value :: sme
			integer(c_int), value :: renorm
			real(c_double) :: eps(*)
		end function
		integer(c_int) function wtb_optprop(ctx, n, ncomp, energy, eps_re, eps_im, out) &
				bind(C, name="wtb_optprop")
			import :: c_int, c_ptr, c_double
			type(c_ptr), value :: ctx
			integer(c_int), value :: n, ncomp
			real(c_double) :: energy(*), eps_re(*), eps_im(*), out(*)
		end function
		integer(c_int) function wtb_exciton_lifetime(ctx, sysdim, ngrid, rlat, n, ncomp, nout, fosc, lft) &
				bind(C, name="wtb_exciton_lifetime")
			import :: c_int, c_ptr, c_double
			type(c_ptr), value :: ctx
			integer(c_int), value :: sysdim, n, ncomp, nout
			integer(c_int) :: ngrid(3)
			real(c_double) :: rlat(*), fosc(*), lft(*)
		end function
		integer(c_int) function wtb_zheev(ctx, want_vectors, n, a, lda, w) bind(C, name="wtb_zheev")
			import :: c_int, c_ptr, c_double
			type(c_ptr), value :: ctx
			integer(c_int), value :: want_vectors, n, lda
			complex(c_double_complex) :: a(*)
			real(c_double) :: w(*)
		end function
	end interface

contains

	! COULOMB_POT string -> WTB_* selector (the cases of matrizelbse, bse_subs.f90:45-91)
	integer(c_int) function wtb_coulomb_id(coultype)
		character(len=*) :: coultype
		select case (trim(adjustl(coultype)))
		case ("V2DK"); wtb_coulomb_id = 1
		case ("V3D"); wtb_coulomb_id = 2
		case ("V3DL"); wtb_coulomb_id = 3
		case ("V2DT"); wtb_coulomb_id = 4
		case ("V2DT2"); wtb_coulomb_id = 5
		case ("V2DOH"); wtb_coulomb_id = 6
		case ("V2DRK"); wtb_coulomb_id = 7
		case ("V1DT"); wtb_coulomb_id = 8
		case ("V0DT"); wtb_coulomb_id = 9
		case default; wtb_coulomb_id = -1
		end select
	end function

	! the arithmetic form of the drop-ins: WTB_GPU_PRECISION=mixed in the environment selects
	! the precision form (COMPLEX*8 matrix and reduction, FP64 tridiagonal solve); unset or
	! any other value keeps the plain FP64 form, the reference's arithmetic.  (The
	! reference's input parser is used unchanged, so the choice cannot be an input key.)
	integer(c_int) function wtb_form_from_env()
		character(len=32) :: v
		integer :: n, st
		wtb_form_from_env = WTB_FP64
		call get_environment_variable("WTB_GPU_PRECISION", v, n, st)
		if (st == 0 .and. n > 0) then
			if (trim(adjustl(v)) == "mixed" .or. trim(adjustl(v)) == "MIXED") wtb_form_from_env = WTB_MIXED
		end if
	end function

	! map a status to the reference's STOP text and stop, as the reference does
	subroutine wtb_check(rc)
		integer(c_int) :: rc
		character(len=256) :: msg
		character(kind=c_char), pointer :: cm(:)
		integer :: i
		if (rc == WTB_OK) return
		msg = ""
		call c_f_pointer(wtb_last_error(), cm, [256])
		do i = 1, 256
			if (cm(i) == c_null_char) exit
			msg(i:i) = cm(i)
		end do
		select case (rc)
		case (4)
			write(*,*) "Wrong Coulomb Potential"
		case (5)
			write(*,*) "Number of valence states higher then ocuppied states"
		case (6)
			write(*,*) "Number of conduction states higher then unocuppied states"
		case (8)
			write(*,*) 'The algorithm failed to compute eigenvalues.'
		case default
			write(*,*) "wtb_gpu error ", rc, ": ", trim(msg)
		end select
		stop
	end subroutine

end module wtb_gpu
