! Test infrastructure only (part of the oracle): a driver around the
! reference's OWN subroutines, compiled from the reference sources where they
! lie (see oracle/Makefile; the include path is given with -I).  It runs the
! BSE pipeline in the order of bsesolver (subprograms/bse_solver-tool-diel.f90:
! 223-620) and bsedielraw (subprograms/bse_diel-tool.f90:60-218) and writes
! every intermediate at full precision to a binary stream file, so that the
! project's restatement and its GPU path can be pinned to the reference's
! arithmetic.  It also times each stage (omp_get_wtime), which serves as the
! reference CPU baseline.  Nothing in the product loads or runs this program.
!
! Input: the reference's own stdin "KEY= value" format (input_read,
! subroutines/module_input_read.f90:235-663).  Harness options come from the
! environment:
!   WTB_DUMP       path of the binary dump (default: no dump)
!   WTB_DUMP_HBSE  1 = also dump the upper triangle of the BSE matrix
!   WTB_SKIP_DIAG  1 = stop after the BSE build (timing of stages a..c only)
! With BSET= T in the input it runs bsesolvertemp's variant instead (subprograms/
! bse_solver-tool-diel-temp.f90): the gap correction tcor added to the scissor,
! f_v - f_c per BSE state, matrizelbsetemp and dielbseptemp.

INCLUDE "subroutines/berry_curvature_subs.f90"
INCLUDE "subroutines/bse_subs.f90"
INCLUDE "subroutines/bse_subs_kpath.f90"
INCLUDE "subroutines/coulomb_pot.f90"
INCLUDE "subroutines/diel-pp-subs.f90"
INCLUDE "subroutines/dos_subs.f90"
INCLUDE "subroutines/general_subs.f90"
INCLUDE "subroutines/mhkpack_subs.f90"
INCLUDE "subroutines/module_input_read.f90"
INCLUDE "subroutines/optics.f90"
INCLUDE "subroutines/hamiltonian_tb.f90"
INCLUDE "subroutines/bse_subs_temp.f90"
INCLUDE "subroutines/pce-subs.f90"

module dumpio
	implicit none
	integer :: du = 77
	logical :: dumping = .false.
contains
	subroutine dump_open(path)
		character(len=*) :: path
		open(unit=du, file=path, access='stream', form='unformatted', status='replace')
		dumping = .true.
	end subroutine
	subroutine dump_r8(tag, n, a)
		character(len=*) :: tag
		integer :: n
		double precision, dimension(n) :: a
		character(len=8) :: t
		if (.not. dumping) return
		t = tag
		write(du) t, int(n,8), int(1,4), a
	end subroutine
	subroutine dump_c16(tag, n, a)
		character(len=*) :: tag
		integer :: n
		double complex, dimension(n) :: a
		character(len=8) :: t
		if (.not. dumping) return
		t = tag
		write(du) t, int(n,8), int(2,4), a
	end subroutine
	subroutine dump_i4(tag, n, a)
		character(len=*) :: tag
		integer :: n
		integer, dimension(n) :: a
		character(len=8) :: t
		if (.not. dumping) return
		t = tag
		write(du) t, int(n,8), int(3,4), a
	end subroutine
end module dumpio

program ref_dump
	use omp_lib
	use input_variables
	use hamiltonian_input_variables
	use dumpio
	implicit none

	integer :: ngkpt, dimbse, i, j, l, h, nw, info, lwork, lwmax, nom
	double precision, allocatable :: kpt(:,:), eigv(:,:), eaux(:), eall(:,:)
	double complex, allocatable :: vector(:,:,:), vaux(:,:)
	integer, allocatable :: nocpk(:), stt(:,:)
	double complex, allocatable :: hrx(:), hry(:), hrz(:), hrsp(:), hrsm(:)
	double complex, allocatable :: hbse(:,:), work(:), tri(:)
	double precision, allocatable :: wbse(:), rwork(:)
	double precision, allocatable :: act(:,:), dielf(:,:,:), actq(:,:), wq(:), dielq(:,:,:)
	double precision :: egap, vc, elux, rr, ii, spinf
	double precision :: t0, t1, tab, tdip, tbuild, tdiag, tosc, teps
	double complex :: matrizelbse, matrizelbsetemp
	double precision :: tcor, ecx, evx, fermidisteh, gapcortemp, gapcortemp2
	double precision, allocatable :: fdeh(:)
	character(len=256) :: envv
	character(len=120) :: line
	logical :: dump_hbse, skip_diag
	integer :: comp, rtype
	rlat = 0.0d0

	call input_read

	call get_environment_variable("WTB_DUMP", envv)
	if (len_trim(envv) > 0) call dump_open(trim(envv))
	call get_environment_variable("WTB_DUMP_HBSE", envv)
	dump_hbse = (trim(envv) == "1")
	call get_environment_variable("WTB_SKIP_DIAG", envv)
	skip_diag = (trim(envv) == "1")

	call OMP_SET_NUM_THREADS(nthreads)
	call hamiltonian_input_read(200, params)

	! mesh sizes as in program main (wtb_main.f90:88-102)
	if (meshtype .eq. "RK3D") then
		call rkmesh(rk, rlat, ngrid)
	else if (meshtype .eq. "RK2D") then
		call rkmesh2D(rk, rlat, ngrid)
	end if

	nw = w90basis
	ngkpt = ngrid(1)*ngrid(2)*ngrid(3)
	dimbse = ngkpt*nc*nv

	call dump_i4("ngrid", 3, ngrid)
	call dump_r8("rlat", 9, rlat)

	allocate(kpt(ngkpt,3))
	call monhkhorst_pack(ngrid(1),ngrid(2),ngrid(3),mshift,rlat(1,:),rlat(2,:),rlat(3,:),kpt)
	call dump_r8("kpt", ngkpt*3, kpt)

	allocate(eaux(nw), vaux(nw,nw), eall(nw,ngkpt))
	allocate(eigv(ngkpt,nc+nv), vector(ngkpt,nc+nv,nw), nocpk(ngkpt))
	allocate(stt(dimbse,4))
	allocate(hrx(dimbse),hry(dimbse),hrz(dimbse),hrsp(dimbse),hrsm(dimbse))

	! finite temperature: the gap correction (bse_solver-tool-diel-temp.f90:262-283)
	tcor = 0.0
	if (bset) then
		select case (ta)
		case("FA")
			tcor = 0.0
		case("VE")
			tcor = gapcortemp(st,phavg,temp)
		case("BE")
			tcor = gapcortemp2(st,phavg,temp)
		case default
			tcor = 0.00
		end select
	end if
	call dump_r8("tcor", 1, (/tcor/))

	! stages (a)+(b): bse_solver-tool-diel.f90:257-313
	egap = 50.0
	t0 = omp_get_wtime()
	do i=1,ngkpt
		call eigsys(nthreads,scs+tcor,exc,nocpk(i),ffactor,kpt(i,1),kpt(i,2),kpt(i,3),nw,nvec,&
			    rlat,rvec,hopmatrices,ihopmatrices,efermi,eaux,vaux)
		eall(:,i) = eaux
		do j=1,nc+nv
			eigv(i,j) = eaux(nocpk(i)-nv+j)
		end do
		if ((eigv(i,nv+1)-eigv(i,nv)) .le. egap) egap = eigv(i,nv+1)-eigv(i,nv)
		do l=1,nc+nv
			do h=1,nw
				vector(i,l,h) = vaux(nocpk(i)-nv+l,h)
			end do
		end do
	end do
	t1 = omp_get_wtime(); tab = t1-t0
	call dump_r8("eall", nw*ngkpt, eall)
	call dump_i4("nocpk", ngkpt, nocpk)
	call dump_r8("eigv", ngkpt*(nc+nv), eigv)
	call dump_c16("vector", ngkpt*(nc+nv)*nw, vector)
	call dump_r8("egap", 1, (/egap/))

	call quantumnumbers2(nw,ngkpt,nc,nv,nocpk,nocpk,stt)
	call dump_i4("stt", dimbse*4, stt)

	! dipoles: bse_solver-tool-diel.f90:365-396
	t0 = omp_get_wtime()
	do i=1,dimbse
		call optsp(eigv(stt(i,4),stt(i,2)),vector(stt(i,4),stt(i,2),:),&
		     eigv(stt(i,4),stt(i,3)),vector(stt(i,4),stt(i,3),:),&
		     kpt(stt(i,4),1),kpt(stt(i,4),2),kpt(stt(i,4),3),ffactor,cshift,&
		     nw,nvec,rlat,rvec,hopmatrices,ihopmatrices,hrx(i),hry(i),hrz(i))
		hrsp(i) = (hrx(i)+cmplx(0.,1.)*hry(i))*(1.0/sqrt(2.))
		hrsm(i) = (hrx(i)-cmplx(0.,1.)*hry(i))*(1.0/sqrt(2.))
	end do
	t1 = omp_get_wtime(); tdip = t1-t0
	call dump_c16("hrx", dimbse, hrx)
	call dump_c16("hry", dimbse, hry)
	call dump_c16("hrz", dimbse, hrz)
	call dump_c16("hrsp", dimbse, hrsp)
	call dump_c16("hrsm", dimbse, hrsm)
	if (bset) then
		! bse_solver-tool-diel-temp.f90:394-399 (fermidisteh sets temp = 1E-8 when it is 0)
		allocate(fdeh(dimbse))
		do i=1,dimbse
			ecx = eigv(stt(i,4),stt(i,3))
			evx = eigv(stt(i,4),stt(i,2))
			fdeh(i) = fermidisteh(ecx,evx,temp)
		end do
		call dump_r8("fdeh", dimbse, fdeh)
	end if

	! stage (c): bse_solver-tool-diel.f90:429-458
	allocate(hbse(dimbse,dimbse), wbse(dimbse))
	t0 = omp_get_wtime()
	hbse = 0.0
	if (bset) then
	!$omp parallel do
	do i=1,dimbse
		do j=i,dimbse
  hbse(i,j)= matrizelbsetemp(coultype,ktol,nw,ediel,lc,ez,w,r0,ngrid,rlat,stt(i,:),eigv(stt(i,4)&
  	    ,stt(i,3)),eigv(stt(i,4),stt(i,2)),vector(stt(i,4)&
            ,stt(i,3),:) ,vector(stt(i,4),stt(i,2),:),kpt(stt(i,4),:),stt(j,:),eigv(stt(j,4),stt(j,3))&
  	    ,eigv(stt(j,4),stt(j,2)) &
            ,vector(stt(j,4),stt(j,3),:),vector(stt(j,4),stt(j,2),:),kpt(stt(j,4),:),temp)
		end do
	end do
	!$omp end parallel do
	else
	!$omp parallel do
	do i=1,dimbse
		do j=i,dimbse
  hbse(i,j)= matrizelbse(coultype,ktol,nw,ediel,lc,ez,w,r0,ngrid,rlat,stt(i,:),eigv(stt(i,4)&
  	    ,stt(i,3)),eigv(stt(i,4),stt(i,2)),vector(stt(i,4)&
            ,stt(i,3),:) ,vector(stt(i,4),stt(i,2),:),kpt(stt(i,4),:),stt(j,:),eigv(stt(j,4),stt(j,3))&
  	    ,eigv(stt(j,4),stt(j,2)) &
            ,vector(stt(j,4),stt(j,3),:),vector(stt(j,4),stt(j,2),:),kpt(stt(j,4),:))
		end do
	end do
	!$omp end parallel do
	end if
	t1 = omp_get_wtime(); tbuild = t1-t0
	if (dump_hbse) then
		allocate(tri(dimbse*(dimbse+1)/2))
		l = 0
		do j=1,dimbse
			do i=1,j
				l = l+1
				tri(l) = hbse(i,j)
			end do
		end do
		call dump_c16("hbse_up", dimbse*(dimbse+1)/2, tri)
		deallocate(tri)
	end if
	! a few rows of the matrix are always dumped (rows 1, 2 and dimbse/2)
	call dump_c16("hrow1", dimbse, hbse(1,:))
	call dump_c16("hrow2", dimbse, hbse(min(2,dimbse),:))
	call dump_c16("hrowm", dimbse, hbse(max(1,dimbse/2),:))

	if (skip_diag) then
		write(*,'(A,I8,A,ES12.5,A,ES12.5,A,ES12.5)') 'TIMES N=',dimbse,' ab=',tab,' dip=',tdip,' build=',tbuild
		stop
	end if

	! stage (d): bse_solver-tool-diel.f90:483-492 (LWORK clipped to 2N-1)
	lwmax = 2*dimbse-1
	allocate(work(lwmax), rwork(3*dimbse-2))
	t0 = omp_get_wtime()
	lwork = -1
	call ZHEEV('Vectors','U',dimbse,hbse,dimbse,wbse,work,lwork,rwork,info)
	lwork = min(lwmax, int(work(1)))
	call ZHEEV('Vectors','U',dimbse,hbse,dimbse,wbse,work,lwork,rwork,info)
	t1 = omp_get_wtime(); tdiag = t1-t0
	if (info > 0) stop "The algorithm failed to compute eigenvalues."
	call dump_r8("wbse", dimbse, wbse)
	call dump_c16("xvec1", dimbse, hbse(:,1))
	if (dump_hbse) call dump_c16("xvecs", dimbse*dimbse, hbse)

	! oscillator strengths: bse_solver-tool-diel.f90:540-602 (DTDIAG=T and CPOL=T)
	allocate(act(dimbse,8))
	t0 = omp_get_wtime()
	if (bset) then
	! bse_solver-tool-diel-temp.f90:570-618
	call dielbseptemp(nthreads,dimbse,hbse,hrx,hrx,fdeh,act(:,1))
	call dielbseptemp(nthreads,dimbse,hbse,hry,hry,fdeh,act(:,2))
	call dielbseptemp(nthreads,dimbse,hbse,hrz,hrz,fdeh,act(:,3))
	call dielbseptemp(nthreads,dimbse,hbse,hrx,hry,fdeh,act(:,4))
	call dielbseptemp(nthreads,dimbse,hbse,hrx,hrz,fdeh,act(:,5))
	call dielbseptemp(nthreads,dimbse,hbse,hry,hrz,fdeh,act(:,6))
	call dielbseptemp(nthreads,dimbse,hbse,hrsp,hrsp,fdeh,act(:,7))
	call dielbseptemp(nthreads,dimbse,hbse,hrsm,hrsm,fdeh,act(:,8))
	else
	call dielbsev(nthreads,dimbse,hbse,hrx,act(:,1))
	call dielbsev(nthreads,dimbse,hbse,hry,act(:,2))
	call dielbsev(nthreads,dimbse,hbse,hrz,act(:,3))
	call dielbsep(nthreads,dimbse,hbse,hrx,hry,act(:,4))
	call dielbsep(nthreads,dimbse,hbse,hrx,hrz,act(:,5))
	call dielbsep(nthreads,dimbse,hbse,hry,hrz,act(:,6))
	call dielbsev(nthreads,dimbse,hbse,hrsp,act(:,7))
	call dielbsev(nthreads,dimbse,hbse,hrsm,act(:,8))
	end if
	t1 = omp_get_wtime(); tosc = t1-t0
	call dump_r8("act", dimbse*8, act)

	! bse_opt_diel.dat round trip in F15.4 (bse_solver-tool-diel.f90:610,
	! bse_diel-tool.f90:117-123)
	allocate(actq(dimbse,6), wq(dimbse))
	do i=1,dimbse
		if (dtfull) then
			write(line,"(7F15.4)") wbse(i),act(i,1),act(i,2),act(i,3),act(i,4),act(i,5),act(i,6)
		else
			write(line,"(7F15.4)") wbse(i),act(i,1),act(i,2),act(i,3),0.0,0.0,0.0
		end if
		read(line,*) wq(i),actq(i,1),actq(i,2),actq(i,3),actq(i,4),actq(i,5),actq(i,6)
	end do
	call dump_r8("wq", dimbse, wq)
	call dump_r8("actq", dimbse*6, actq)

	! stage (e): bsedielraw loop, bse_diel-tool.f90:111,135-201
	call vcell3D(rlat, vc)
	if (systype .eq. "NP") then
		spinf = 2.0
	else
		spinf = 1.0
	end if
	nom = int(numbse)
	allocate(dielf(nom,3,6), dielq(nom,3,6))
	t0 = omp_get_wtime()
	do j=1,nom
		elux = ebse0 + dble(((ebsef-ebse0)*(j-1))/(numbse-1.))
		do comp=1,6
			rtype = 1
			if (comp > 3) rtype = 0
			call optdiel(rtype,vc,dimbse,ngrid,elux,wq,actq(:,comp),cshift,rr,ii)
			dielq(j,1,comp) = elux
			dielq(j,2,comp) = rr
			dielq(j,3,comp) = ii
			call optdiel(rtype,vc,dimbse,ngrid,elux,wbse,act(:,comp),cshift,rr,ii)
			dielf(j,1,comp) = elux
			dielf(j,2,comp) = rr
			dielf(j,3,comp) = ii
		end do
	end do
	if (renorm) then
		do comp=1,6
			call imagrenorm(dimbse,nom,wq,dielq(:,:,comp))
			call imagrenorm(dimbse,nom,wbse,dielf(:,:,comp))
		end do
	end if
	t1 = omp_get_wtime(); teps = (t1-t0)/2.0
	call dump_r8("dielq", nom*3*6, dielq)
	call dump_r8("dielf", nom*3*6, dielf)
	call dump_r8("vc3d", 1, (/vc/))
	call dump_r8("spinf", 1, (/spinf/))

	write(*,'(A,I8,A,I4,A,ES12.5,A,ES12.5,A,ES12.5,A,ES12.5,A,ES12.5,A,ES12.5)') 'TIMES N=',dimbse, &
		' threads=',nthreads,' ab=',tab,' dip=',tdip,' build=',tbuild,' diag=',tdiag,' osc=',tosc,' eps=',teps
	if (dumping) close(du)
end program ref_dump
