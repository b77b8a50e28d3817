! Drop-in replacements for the reference's bsesolver
! (subprograms/bse_solver-tool-diel.f90:2-660), bsesolvertemp
! (subprograms/bse_solver-tool-diel-temp.f90:2-660), bsebnds
! (subprograms/bse_kpath-tool.f90:3-479), bsebndstemp
! (subprograms/bse_kpath-tool-temp.f90), bsedielraw
! (subprograms/bse_diel-tool.f90:5-245), bsedielrawpol
! (subprograms/bse_diel-tool-pol.f90:5-242), spdielraw (subprograms/sp_diel-tool.f90:4-239)
! and spdielrawpol (subprograms/sp_diel-tool-pol.f90:4-235): same names, same argument
! lists, same output files, with the computation done by libwtb_bse.so on the GPU.
!
! Use: INCLUDE this file (after the reference's subroutines/module_input_read.f90
! and wantibexos_amd/fortran/wtb_gpu.f90) in place of those reference files,
! and link -lwtb_bse.  Arguments the GPU path does not use (numdos, kpaths,
! orbw, exc, rk, meshtype, ...) are accepted and ignored as in the reference.

subroutine bsesolver(nthreads,outputfolder,calcparms,ngrid,nc,nv,numdos, &
		     ebse0,ebsef,numbse,sme,ktol,params,kpaths,kpathsbse,orbw,ediel, &
		     exc,mshift,coultype,ez,w1,r0,lc,rk,meshtype,bsewf,excwf0,excwff,&
		     dtfull,cpol,tmcoef)
	implicit none
	integer :: nthreads
	integer,dimension(3) :: ngrid
	integer :: nc,nv
	double precision :: numdos, ebse0, ebsef, numbse, sme, exc, rk, ktol
	double precision,dimension(3) :: mshift, ediel
	character(len=70) :: params, orbw, kpaths, kpathsbse, outputfolder, calcparms, meshtype
	character(len=5) :: coultype
	logical :: bsewf, cpol, dtfull, tmcoef
	integer :: excwf0, excwff
	double precision :: ez, w1, lc, r0
	double precision :: zero

	zero = 0d0
	call wtb_bse_dropin(nthreads,outputfolder,ngrid,nc,nv,ebse0,ebsef,numbse,sme,ktol,params, &
			    ediel,mshift,coultype,ez,w1,r0,lc,bsewf,excwf0,excwff,dtfull,cpol,tmcoef, &
			    .false.,zero,zero,"FA",zero,.false.)
end subroutine bsesolver

! bsesolvertemp (subprograms/bse_solver-tool-diel-temp.f90:2-660): the same with BSET= T.
! Like the reference (fermidisteh, general_subs.f90:620-624), TEMP = 0 is set to 1E-8.
subroutine bsesolvertemp(nthreads,outputfolder,calcparms,ngrid,nc,nv,numdos, &
		     ebse0,ebsef,numbse,sme,ktol,params,kpaths,kpathsbse,orbw,ediel, &
		     exc,mshift,coultype,ez,w1,r0,lc,rk,meshtype,bsewf,excwf0,excwff,&
		     dtfull,cpol,tmcoef,st,phavg,ta,temp)
	implicit none
	integer :: nthreads
	integer,dimension(3) :: ngrid
	integer :: nc,nv
	double precision :: numdos, ebse0, ebsef, numbse, sme, exc, rk, ktol
	double precision,dimension(3) :: mshift, ediel
	character(len=70) :: params, orbw, kpaths, kpathsbse, outputfolder, calcparms, meshtype
	character(len=5) :: coultype
	logical :: bsewf, cpol, dtfull, tmcoef
	integer :: excwf0, excwff
	double precision :: ez, w1, lc, r0
	double precision :: st, phavg, temp
	character(len=2) :: ta

	call wtb_bse_dropin(nthreads,outputfolder,ngrid,nc,nv,ebse0,ebsef,numbse,sme,ktol,params, &
			    ediel,mshift,coultype,ez,w1,r0,lc,bsewf,excwf0,excwff,dtfull,cpol,tmcoef, &
			    .true.,st,phavg,ta,temp,.false.)
	if (temp .eq. 0d0) temp = 1E-8
end subroutine bsesolvertemp

! spoptics (subprograms/sp_solver-tool-diel.f90:3-323), DIEL= T: the single-particle
! table of the BSE drop-in without the BSE solve (sp_optics.dat, sp_optics-pol.dat and,
! with TMCOEF= T, tm_coef*.dat).  The table uses no Coulomb potential; the plan behind
! wtb_sp_table still wants a valid one, so the potential's own parameters are neutral.
subroutine spoptics(nthreads,outputfolder,calcparms,ngrid,nc,nv,numdos, &
		    ebse0,ebsef,numbse,sme,ktol,params,kpaths,kpathsbse,orbw,ediel, &
		    exc,mshift,coultype,rk,meshtype,tmcoef)
	implicit none
	integer :: nthreads
	integer,dimension(3) :: ngrid
	integer :: nc,nv
	double precision :: numdos, ebse0, ebsef, numbse, sme, exc, rk, ktol
	double precision,dimension(3) :: mshift, ediel
	character(len=70) :: params, orbw, kpaths, kpathsbse, outputfolder, calcparms, meshtype
	character(len=5) :: coultype
	logical :: tmcoef
	double precision :: zero, one

	zero = 0d0
	one = 1d0
	call wtb_bse_dropin(nthreads,outputfolder,ngrid,nc,nv,ebse0,ebsef,numbse,sme,ktol,params, &
			    ediel,mshift,coultype,one,zero,one,one,.false.,1,1,.true.,.false.,tmcoef, &
			    .false.,zero,zero,"FA",zero,.true.)
end subroutine spoptics


! The body of both: model and parameters into the C structures, one wtb_bsesolver(_wf)
! call, the reference's output files.
subroutine wtb_bse_dropin(nthreads,outputfolder,ngrid,nc,nv,ebse0,ebsef,numbse,sme,ktol,params, &
			  ediel,mshift,coultype,ez,w1,r0,lc,bsewf,excwf0,excwff,dtfull,cpol,tmcoef, &
			  bset,st,phavg,ta,temp,sponly)
	use iso_c_binding
	use hamiltonian_input_variables
	use wtb_gpu
	implicit none
	integer :: nthreads
	integer,dimension(3) :: ngrid
	integer :: nc,nv
	double precision :: ebse0, ebsef, numbse, sme, ktol
	double precision,dimension(3) :: mshift, ediel
	character(len=70) :: params, outputfolder
	character(len=5) :: coultype
	logical :: bsewf, cpol, dtfull, bset, tmcoef, sponly
	integer :: excwf0, excwff
	double precision :: ez, w1, lc, r0
	double precision :: st, phavg, temp
	character(len=2) :: ta

	type(c_ptr) :: ctx
	real, allocatable :: vecres(:,:)
	type(wtb_tb_model) :: m
	type(wtb_bse_params) :: p
	double precision, allocatable, target :: t_hop(:,:,:), t_ihop(:,:,:)
	integer, allocatable, target :: t_rvec(:,:), t_ff(:)
	double precision, target :: t_rlat(3,3)
	double precision, allocatable :: w(:), act(:,:)
	double complex, allocatable :: wf(:,:)
	double precision, allocatable :: kpt(:,:)
	integer, allocatable :: nocpk(:), stt(:,:)
	double precision :: egap
	integer :: dimbse, i, erro, nwf, ngkpt
	integer(c_int) :: rc

	call hamiltonian_input_read(200, params)

	allocate(t_hop(nvec,w90basis,w90basis), t_ihop(nvec,w90basis,w90basis))
	allocate(t_rvec(nvec,3), t_ff(nvec))
	t_hop = hopmatrices
	t_ihop = ihopmatrices
	t_rvec = rvec
	t_ff = ffactor(1:nvec)
	t_rlat = rlat
	m%nw = w90basis
	m%nvec = nvec
	m%rlat = c_loc(t_rlat)
	m%rvec = c_loc(t_rvec)
	m%ffactor = c_loc(t_ff)
	m%hop = c_loc(t_hop)
	m%ihop = c_loc(t_ihop)
	m%efermi = efermi
	m%scs = scs

	p%ngrid = ngrid
	p%mshift = mshift
	p%nc = nc
	p%nv = nv
	p%coultype = wtb_coulomb_id(coultype)
	p%ediel = ediel
	p%lc = lc
	p%ez = ez
	p%w = w1
	p%r0 = r0
	p%ktol = ktol
	p%cshift = sme
	p%ebse0 = ebse0
	p%ebsef = ebsef
	p%numbse = int(numbse)
	p%dtfull = merge(1, 0, dtfull)
	p%cpol = merge(1, 0, cpol)
	p%renorm = 1
	p%f154 = 1
	p%precision = wtb_form_from_env()
	p%bset = merge(1, 0, bset)
	p%ta = WTB_TA_FA
	if (ta .eq. "VE") p%ta = WTB_TA_VE
	if (ta .eq. "BE") p%ta = WTB_TA_BE
	p%temp = temp
	p%st = st
	p%phavg = phavg

	dimbse = ngrid(1)*ngrid(2)*ngrid(3)*nc*nv
	allocate(w(dimbse), act(dimbse,8))

	rc = wtb_create(0_c_int, ctx)
	call wtb_check(rc)
	if (sponly) then
		continue
	else if (bsewf) then
		! BSE_WF: the vectors of excitons excwf0 .. excwff and what excwf2 prints beside them
		if (excwf0 < 1 .or. excwff > dimbse .or. excwff < excwf0) stop "Wrong BSE_WF exciton range"
		nwf = excwff - excwf0 + 1
		ngkpt = ngrid(1)*ngrid(2)*ngrid(3)
		allocate(wf(dimbse,nwf), kpt(ngkpt,3), nocpk(ngkpt), stt(dimbse,4))
		rc = wtb_bsesolver_wf(ctx, m, p, w, act, c_null_ptr, egap, int(excwf0-1,c_int), int(nwf,c_int), &
				      wf, kpt, nocpk)
	else
		rc = wtb_bsesolver(ctx, m, p, w, act, c_null_ptr, egap)
	end if
	call wtb_check(rc)
	! the single-particle side output (bse_solver-tool-diel.f90:336-420), which is all
	! that spoptics (subprograms/sp_solver-tool-diel.f90) writes
	allocate(vecres(dimbse,15))
	rc = wtb_sp_table(ctx, m, p, vecres)
	call wtb_check(rc)
	rc = wtb_destroy(ctx)
	OPEN(UNIT=401, FILE=trim(outputfolder)//"sp_optics.dat",STATUS='unknown', IOSTAT=erro)
	if (erro/=0) stop "Error opening bse_opt_diel-pol output file"
	OPEN(UNIT=403, FILE=trim(outputfolder)//"sp_optics-pol.dat",STATUS='unknown', IOSTAT=erro)
	if (erro/=0) stop "Error opening sp_optics-pol output file"
	write(401,*) "#","  ", "energy","  ","xx","  ","yy","  ","zz","  ","xy","  ","xz","  ","yz"
	write(403,*) "#","  ", "energy","  ","xx","  ","yy","  ","zz","  ","sp","  ","sm"
	if (tmcoef) then
		OPEN(UNIT=402, FILE=trim(outputfolder)//"tm_coef.dat",STATUS='unknown', IOSTAT=erro)
		if (erro/=0) stop "Error opening tm_coef output file"
		OPEN(UNIT=404, FILE=trim(outputfolder)//"tm_coef-pol.dat",STATUS='unknown', IOSTAT=erro)
		if (erro/=0) stop "Error opening tm_coef-pol output file"
		write(402,*) "number of kpoints:",ngrid(1)*ngrid(2)*ngrid(3)
		write(402,*) "number of conduction states",nc
		write(402,*) "number of valence states",nv
		write(402,*) "#"," ", "kx", " ", "ky"," ", "kz"," ","nocp"," ", "nc"," ", "nv","  ", "energy","  ","xx","  ",&
			      "yy","  ","zz","  ","xy","  ","xz","  ","yz"
		write(404,*) "number of kpoints:",ngrid(1)*ngrid(2)*ngrid(3)
		write(404,*) "number of conduction states",nc
		write(404,*) "number of valence states",nv
		write(404,*) "#"," ", "kx", " ", "ky"," ", "kz"," ","nocp"," ", "nc"," ", "nv","  ", "energy","  ","xx","  ",&
			      "yy","  ","zz","  ","sp","  ","sm"
	end if
	do i=1,dimbse
		write(401,"(7F15.4)") vecres(i,7),vecres(i,8),vecres(i,9),vecres(i,10),vecres(i,11),vecres(i,12),vecres(i,13)
		write(403,"(6F15.4)") vecres(i,7),vecres(i,8),vecres(i,9),vecres(i,10),vecres(i,14),vecres(i,15)
		if (tmcoef) then
			write(402,"(3F10.4,3I10.0,7F10.4)") vecres(i,1),vecres(i,2),vecres(i,3),int(vecres(i,4)), &
				int(vecres(i,5)),int(vecres(i,6)),vecres(i,7),vecres(i,8),vecres(i,9),vecres(i,10), &
				vecres(i,11),vecres(i,12),vecres(i,13)
			write(404,"(3F10.4,3I10.0,6F10.4)") vecres(i,1),vecres(i,2),vecres(i,3),int(vecres(i,4)), &
				int(vecres(i,5)),int(vecres(i,6)),vecres(i,7),vecres(i,8),vecres(i,9),vecres(i,10), &
				vecres(i,14),vecres(i,15)
		end if
	end do
	close(401)
	close(403)
	if (tmcoef) then
		close(402)
		close(404)
	end if
	deallocate(vecres)
	if (sponly) then
		close(200)
		deallocate(w, act)
		deallocate(rvec,hopmatrices,ihopmatrices,ffactor)
		return
	end if

	OPEN(UNIT=300, FILE=trim(outputfolder)//"log_bse-diel.dat",STATUS='unknown', IOSTAT=erro)
	if (erro/=0) stop "Error opening log_bse-diel output file"
	OPEN(UNIT=301, FILE=trim(outputfolder)//"bse_opt_diel.dat",STATUS='unknown', IOSTAT=erro)
	if (erro/=0) stop "Error opening bse_opt_diel output file"
	OPEN(UNIT=302, FILE=trim(outputfolder)//"bse_opt_diel-pol.dat",STATUS='unknown', IOSTAT=erro)
	if (erro/=0) stop "Error opening output file"

	write(300,*) 'threads:', nthreads
	if (bset) write(300,*) 'temperature', temp
	write(300,*) 'grid:',ngrid(1),ngrid(2),ngrid(3)
	write(300,"(A14,1E15.4)") 'ktol-coulomb:', ktol
	write(300,"(A13,3F15.4)") 'kmesh shift:',mshift(1),mshift(2),mshift(3)
	write(300,*) 'conduction bands, nc:',nc,'  ','valence bands, nv:',nv
	write(300,*) 'direct gap:', egap
	write(300,*) 'exciton ground state',w(1)
	write(300,*) 'computed on the GPU (libwtb_bse)'

	write(301,*) "#","  ", "exciton energy","  ","xx","  ","yy","  ","zz"," ","xy","  ","xz","  ","yz"
	write(302,*) "#","  ", "exciton energy","  ","xx","  ","yy","  ","zz"," ","sp","  ","sm"
	do i=1,dimbse
		write(301,"(7F15.4)") w(i),act(i,1),act(i,2),act(i,3),act(i,4),act(i,5),act(i,6)
		if (cpol) write(302,"(6F15.4)") w(i),act(i,1),act(i,2),act(i,3),act(i,7),act(i,8)
	end do
	close(300)
	close(301)
	close(302)
	close(200)
	if (bsewf) then
		call quantumnumbers2(w90basis,ngkpt,nc,nv,nocpk,nocpk,stt)
		do i=excwf0,excwff
			call excwf2(outputfolder,ngkpt,kpt,nc,nv,nocpk,stt,w(i),i,wf(:,i-excwf0+1))
		end do
		deallocate(wf, kpt, nocpk, stt)
	end if
	deallocate(rvec,hopmatrices,ihopmatrices,ffactor)
end subroutine wtb_bse_dropin


! bsedielraw (subprograms/bse_diel-tool.f90:5-245) and spdielraw
! (subprograms/sp_diel-tool.f90:4-239) differ only in their file names: both read
! E, xx, yy, zz, xy, xz, yz per transition and write eps(omega) of the six components
! (optdiel's diagonal form for xx, yy, zz, the off-diagonal one for the others), with
! the same imagrenorm.  wtb_bse_spectrum computes the six spectra on the GPU.
subroutine bsedielraw(nthreads,outputfolder,renorm,params,ngrid,nc,nv,ebse0,ebsef,numbse,sme)
	implicit none
	integer :: nthreads, nc, nv
	integer,dimension(3) :: ngrid
	character(len=70) :: outputfolder, params
	logical :: renorm
	double precision :: ebse0, ebsef, numbse, sme
	call wtb_dielraw_files(outputfolder,renorm,params,ngrid,nc,nv,ebse0,ebsef,numbse,sme, &
			       "bse_opt_diel.dat", "bse_diel_")
end subroutine bsedielraw

subroutine spdielraw(nthreads,outputfolder,renorm,params,ngrid,nc,nv,ebse0,ebsef,numbse,sme)
	implicit none
	integer :: nthreads, nc, nv
	integer,dimension(3) :: ngrid
	character(len=70) :: outputfolder, params
	logical :: renorm
	double precision :: ebse0, ebsef, numbse, sme
	call wtb_dielraw_files(outputfolder,renorm,params,ngrid,nc,nv,ebse0,ebsef,numbse,sme, &
			       "sp_optics.dat", "sp_diel_")
end subroutine spdielraw

subroutine wtb_dielraw_files(outputfolder,renorm,params,ngrid,nc,nv,ebse0,ebsef,numbse,sme,fin,fout)
	use iso_c_binding
	use hamiltonian_input_variables
	use wtb_gpu
	implicit none
	integer :: nc, nv
	integer,dimension(3) :: ngrid
	character(len=70) :: outputfolder, params
	character(len=*) :: fin, fout
	logical :: renorm
	double precision :: ebse0, ebsef, numbse, sme

	type(c_ptr) :: ctx
	double precision, allocatable :: ebse(:), act6(:,:), eps(:,:,:)
	double precision :: spinf
	character(len=2) :: aread
	integer :: dimbse, j, nom, erro, c
	integer(c_int) :: rc
	integer, parameter :: order(6) = (/1, 4, 5, 2, 6, 3/)   ! files xx,xy,xz,yy,yz,zz
	character(len=2), parameter :: tag(6) = (/"xx","yy","zz","xy","xz","yz"/)

	call hamiltonian_input_read(200, params)
	spinf = 1.0
	if (systype .eq. "NP") spinf = 2.0

	dimbse = ngrid(1)*ngrid(2)*ngrid(3)*nc*nv
	nom = int(numbse)
	allocate(ebse(dimbse), act6(dimbse,6), eps(nom,2,6))
	OPEN(UNIT=100, FILE=trim(outputfolder)//fin,STATUS='unknown', IOSTAT=erro)
	if (erro/=0) stop "Error opening "//fin//" input file"
	read(100,*) aread
	do j=1,dimbse
		read(100,*) ebse(j),act6(j,1),act6(j,2),act6(j,3),act6(j,4),act6(j,5),act6(j,6)
	end do
	close(100)

	rc = wtb_create(0_c_int, ctx)
	call wtb_check(rc)
	rc = wtb_bse_spectrum(ctx, dimbse, ebse, act6, rlat, ngrid, ebse0, ebsef, nom, sme, &
			      merge(1, 0, renorm), eps)
	call wtb_check(rc)
	rc = wtb_destroy(ctx)

	do c=1,6
		OPEN(UNIT=300+c, FILE=trim(outputfolder)//fout//tag(order(c))//".dat", &
		     STATUS='unknown', IOSTAT=erro)
		if (erro/=0) stop "Error opening "//fout//" output file"
		write(300+c,*) "#","  ","energy","  ","real","  ","imag"
		do j=1,nom
			write(300+c,*) ebse0 + dble(((ebsef-ebse0)*(j-1))/(numbse-1.)), &
				real(spinf*eps(j,1,order(c))), real(spinf*eps(j,2,order(c)))
		end do
		close(300+c)
	end do
	close(200)
	deallocate(ebse, act6, eps)
	deallocate(rvec,hopmatrices,ihopmatrices,ffactor)
end subroutine wtb_dielraw_files


! bsebnds (subprograms/bse_kpath-tool.f90:3-479): exciton bands along the KPATH_BSE path.
! The path comes from the reference's own kpathbse (which also writes KLABELS-BSE.dat);
! every q-point's matrix is built and diagonalised on the GPU (wtb_bsebnds, q-points on
! four streams); bands_bse.dat is written as the reference writes it (:437-446).
subroutine bsebnds(nthreads,outputfolder,calcparms,ngrid,nc,nv,numdos, &
		     ebse0,ebsef,numbse,sme,ktol,params,kpaths,kpathsbse,orbw,ediel, &
		     exc,mshift,coultype,ez,w1,r0,lc,rk,meshtype,bsewf,excwf0,excwff)
	implicit none
	integer :: nthreads
	integer,dimension(3) :: ngrid
	integer :: nc,nv
	double precision :: numdos, ebse0, ebsef, numbse, sme, exc, rk, ktol
	double precision,dimension(3) :: mshift, ediel
	character(len=70) :: params, orbw, kpaths, kpathsbse, outputfolder, calcparms, meshtype
	character(len=5) :: coultype
	logical :: bsewf
	integer :: excwf0, excwff
	double precision :: ez, w1, lc, r0
	double precision :: zero

	zero = 0d0
	call wtb_bands_dropin(nthreads,outputfolder,ngrid,nc,nv,ebse0,ebsef,numbse,sme,ktol,params, &
			      kpathsbse,ediel,mshift,coultype,ez,w1,r0,lc,bsewf,excwf0,excwff, &
			      .false.,zero,zero,"FA",zero)
end subroutine bsebnds

! bsebndstemp (subprograms/bse_kpath-tool-temp.f90): the same at finite temperature.
subroutine bsebndstemp(nthreads,outputfolder,calcparms,ngrid,nc,nv,numdos, &
		     ebse0,ebsef,numbse,sme,ktol,params,kpaths,kpathsbse,orbw,ediel, &
		     exc,mshift,coultype,ez,w1,r0,lc,rk,meshtype,bsewf,excwf0,excwff,&
		     st,phavg,ta,temp)
	implicit none
	integer :: nthreads
	integer,dimension(3) :: ngrid
	integer :: nc,nv
	double precision :: numdos, ebse0, ebsef, numbse, sme, exc, rk, ktol
	double precision,dimension(3) :: mshift, ediel
	character(len=70) :: params, orbw, kpaths, kpathsbse, outputfolder, calcparms, meshtype
	character(len=5) :: coultype
	logical :: bsewf
	integer :: excwf0, excwff
	double precision :: ez, w1, lc, r0
	double precision :: st, phavg, temp
	character(len=2) :: ta

	call wtb_bands_dropin(nthreads,outputfolder,ngrid,nc,nv,ebse0,ebsef,numbse,sme,ktol,params, &
			      kpathsbse,ediel,mshift,coultype,ez,w1,r0,lc,bsewf,excwf0,excwff, &
			      .true.,st,phavg,ta,temp)
	if (temp .eq. 0d0) temp = 1E-8
end subroutine bsebndstemp

subroutine wtb_bands_dropin(nthreads,outputfolder,ngrid,nc,nv,ebse0,ebsef,numbse,sme,ktol,params, &
			    kpathsbse,ediel,mshift,coultype,ez,w1,r0,lc,bsewf,excwf0,excwff,bset,st,phavg,ta,temp)
	use iso_c_binding
	use hamiltonian_input_variables
	use wtb_gpu
	implicit none
	integer :: nthreads
	integer,dimension(3) :: ngrid
	integer :: nc,nv
	double precision :: ebse0, ebsef, numbse, sme, ktol
	double precision,dimension(3) :: mshift, ediel
	character(len=70) :: params, kpathsbse, outputfolder
	character(len=5) :: coultype
	logical :: bsewf, bset
	integer :: excwf0, excwff
	double precision :: ez, w1, lc, r0
	double precision :: st, phavg, temp
	character(len=2) :: ta

	type(c_ptr) :: ctx
	type(wtb_tb_model) :: m
	type(wtb_bse_params) :: p
	double precision, allocatable, target :: t_hop(:,:,:), t_ihop(:,:,:)
	integer, allocatable, target :: t_rvec(:,:), t_ff(:)
	double precision, target :: t_rlat(3,3)
	double precision, allocatable :: ks(:,:), qauxv(:,:), exk(:,:), kpt(:,:)
	double complex, allocatable :: wf(:,:,:)
	integer, allocatable :: nocpk(:), stt(:,:)
	integer :: nks, nkpts, nq, dimbse, i, i2, erro, nwf, ngkpt
	integer(c_int) :: rc

	OPEN(UNIT=500, FILE= kpathsbse,STATUS='old', IOSTAT=erro)
	if (erro/=0) stop "Error opening bse-kpath input file"
	OPEN(UNIT=300, FILE=trim(outputfolder)//"log_bse_kpath.dat",STATUS='unknown', IOSTAT=erro)
	if (erro/=0) stop "Error opening log_bse_kpath output file"
	OPEN(UNIT=400, FILE=trim(outputfolder)//"bands_bse.dat",STATUS='unknown', IOSTAT=erro)
	if (erro/=0) stop "Error opening bands_bse output file"

	call hamiltonian_input_read(200, params)
	read(500,*) nks
	read(500,*) nkpts
	allocate(ks(nks,3))
	do i=1,nks
		read(500,*) ks(i,1),ks(i,2),ks(i,3)
	end do
	nq = (nks/2)*nkpts
	allocate(qauxv(nq,4))
	call kpathbse(outputfolder,rlat(1,:),rlat(2,:),rlat(3,:),nks,ks,nkpts,qauxv)

	allocate(t_hop(nvec,w90basis,w90basis), t_ihop(nvec,w90basis,w90basis))
	allocate(t_rvec(nvec,3), t_ff(nvec))
	t_hop = hopmatrices
	t_ihop = ihopmatrices
	t_rvec = rvec
	t_ff = ffactor(1:nvec)
	t_rlat = rlat
	m%nw = w90basis
	m%nvec = nvec
	m%rlat = c_loc(t_rlat)
	m%rvec = c_loc(t_rvec)
	m%ffactor = c_loc(t_ff)
	m%hop = c_loc(t_hop)
	m%ihop = c_loc(t_ihop)
	m%efermi = efermi
	m%scs = scs

	p%ngrid = ngrid
	p%mshift = mshift
	p%nc = nc
	p%nv = nv
	p%coultype = wtb_coulomb_id(coultype)
	p%ediel = ediel
	p%lc = lc
	p%ez = ez
	p%w = w1
	p%r0 = r0
	p%ktol = ktol
	p%cshift = sme
	p%ebse0 = ebse0
	p%ebsef = ebsef
	p%numbse = max(2, int(numbse))
	p%dtfull = 0
	p%cpol = 0
	p%renorm = 0
	p%f154 = 0
	p%precision = wtb_form_from_env()
	p%bset = merge(1, 0, bset)
	p%ta = WTB_TA_FA
	if (ta .eq. "VE") p%ta = WTB_TA_VE
	if (ta .eq. "BE") p%ta = WTB_TA_BE
	p%temp = temp
	p%st = st
	p%phavg = phavg

	dimbse = ngrid(1)*ngrid(2)*ngrid(3)*nc*nv
	allocate(exk(dimbse,nq))
	write(300,*) 'threads:', nthreads
	if (bset) write(300,*) 'temperature', temp
	write(300,*) 'grid:',ngrid(1),ngrid(2),ngrid(3)
	write(300,*) 'ktol-coulomb:', ktol
	write(300,*) 'kmesh shift:',mshift(1),mshift(2),mshift(3)
	write(300,*) 'conduction bands, nc:',nc,'  ','valence bands, nv:',nv
	write(300,*) 'number of kpoints in the path:','   ',nq
	write(300,*) 'computed on the GPU (libwtb_bse)'

	rc = wtb_create(0_c_int, ctx)
	call wtb_check(rc)
	if (bsewf) then
		! BSE_WF: the vectors of excitons excwf0 .. excwff at every q, and what excwfi2
		! prints beside them (bse_kpath-tool.f90:402-410)
		if (excwf0 < 1 .or. excwff > dimbse .or. excwff < excwf0) stop "Wrong BSE_WF exciton range"
		nwf = excwff - excwf0 + 1
		ngkpt = ngrid(1)*ngrid(2)*ngrid(3)
		allocate(wf(dimbse,nwf,nq), kpt(ngkpt,3), nocpk(ngkpt), stt(dimbse,4))
		rc = wtb_bsebnds_wf(ctx, m, p, int(nq,c_int), qauxv, 4_c_int, exk, int(excwf0-1,c_int), &
				    int(nwf,c_int), wf, kpt, nocpk)
	else
		rc = wtb_bsebnds(ctx, m, p, int(nq,c_int), qauxv, 4_c_int, exk)
	end if
	call wtb_check(rc)
	rc = wtb_destroy(ctx)
	if (bsewf) then
		call quantumnumbers2(w90basis,ngkpt,nc,nv,nocpk,nocpk,stt)
		do i=1,nq
			do i2=excwf0,excwff
				call excwfi2(outputfolder,ngkpt,kpt,qauxv(i,:),nc,nv,nocpk,stt,exk(i2,i),i2,i, &
					     wf(:,i2-excwf0+1,i))
			end do
		end do
		deallocate(wf, kpt, nocpk, stt)
	end if

	do i2=1,dimbse
		do i=1,nq
			write(400,*) real(qauxv(i,1)),exk(i2,i)
		end do
		write(400,*)
	end do

	close(200)
	close(300)
	close(400)
	close(500)
	deallocate(ks, qauxv, exk)
	deallocate(rvec,hopmatrices,ihopmatrices,ffactor)
end subroutine wtb_bands_dropin


! bsedielrawpol (subprograms/bse_diel-tool-pol.f90:5-242) and spdielrawpol
! (subprograms/sp_diel-tool-pol.f90:4-235): eps(omega) for the x, y, z, sigma+ and
! sigma- polarisations from bse_opt_diel-pol.dat or sp_optics-pol.dat, every one with
! the diagonal (delta = 1) form of optdiel.  wtb_bse_spectrum gives that form to its
! first three components, so the five spectra take two calls.
subroutine bsedielrawpol(nthreads,outputfolder,renorm,params,ngrid,nc,nv,ebse0,ebsef,numbse,sme)
	implicit none
	integer :: nthreads, nc, nv
	integer,dimension(3) :: ngrid
	character(len=70) :: outputfolder, params
	logical :: renorm
	double precision :: ebse0, ebsef, numbse, sme
	call wtb_dielrawpol_files(outputfolder,renorm,params,ngrid,nc,nv,ebse0,ebsef,numbse,sme, &
				  "bse_opt_diel-pol.dat", "bse_diel-pol_")
end subroutine bsedielrawpol

subroutine spdielrawpol(nthreads,outputfolder,renorm,params,ngrid,nc,nv,ebse0,ebsef,numbse,sme)
	implicit none
	integer :: nthreads, nc, nv
	integer,dimension(3) :: ngrid
	character(len=70) :: outputfolder, params
	logical :: renorm
	double precision :: ebse0, ebsef, numbse, sme
	call wtb_dielrawpol_files(outputfolder,renorm,params,ngrid,nc,nv,ebse0,ebsef,numbse,sme, &
				  "sp_optics-pol.dat", "sp_diel-pol_")
end subroutine spdielrawpol

subroutine wtb_dielrawpol_files(outputfolder,renorm,params,ngrid,nc,nv,ebse0,ebsef,numbse,sme,fin,fout)
	use iso_c_binding
	use hamiltonian_input_variables
	use wtb_gpu
	implicit none
	integer :: nc, nv
	integer,dimension(3) :: ngrid
	character(len=70) :: outputfolder, params
	character(len=*) :: fin, fout
	logical :: renorm
	double precision :: ebse0, ebsef, numbse, sme

	type(c_ptr) :: ctx
	double precision, allocatable :: ebse(:), a1(:,:), a2(:,:), eps1(:,:,:), eps2(:,:,:)
	double precision :: spinf, elux
	character(len=2) :: aread
	integer :: dimbse, j, nom, erro, c
	integer(c_int) :: rc
	character(len=2), parameter :: tag(5) = (/"x ","y ","z ","sp","sm"/)

	call hamiltonian_input_read(200, params)
	spinf = 1.0
	if (systype .eq. "NP") spinf = 2.0

	dimbse = ngrid(1)*ngrid(2)*ngrid(3)*nc*nv
	nom = int(numbse)
	allocate(ebse(dimbse), a1(dimbse,6), a2(dimbse,6), eps1(nom,2,6), eps2(nom,2,6))
	a1 = 0d0
	a2 = 0d0
	OPEN(UNIT=100, FILE=trim(outputfolder)//fin,STATUS='unknown', IOSTAT=erro)
	if (erro/=0) stop "Error opening "//fin//" input file"
	read(100,*) aread
	do j=1,dimbse
		read(100,*) ebse(j),a1(j,1),a1(j,2),a1(j,3),a2(j,1),a2(j,2)
	end do
	close(100)

	rc = wtb_create(0_c_int, ctx)
	call wtb_check(rc)
	rc = wtb_bse_spectrum(ctx, dimbse, ebse, a1, rlat, ngrid, ebse0, ebsef, nom, sme, &
			      merge(1, 0, renorm), eps1)
	call wtb_check(rc)
	rc = wtb_bse_spectrum(ctx, dimbse, ebse, a2, rlat, ngrid, ebse0, ebsef, nom, sme, &
			      merge(1, 0, renorm), eps2)
	call wtb_check(rc)
	rc = wtb_destroy(ctx)

	do c=1,5
		OPEN(UNIT=300+c, FILE=trim(outputfolder)//fout//trim(tag(c))//".dat", &
		     STATUS='unknown', IOSTAT=erro)
		if (erro/=0) stop "Error opening "//fout//" output file"
		write(300+c,*) "#","  ","energy","  ","real","  ","imag"
		do j=1,nom
			elux = ebse0 + dble(((ebsef-ebse0)*(j-1))/(numbse-1.))
			if (c <= 3) then
				write(300+c,*) elux, real(spinf*eps1(j,1,c)), real(spinf*eps1(j,2,c))
			else
				write(300+c,*) elux, real(spinf*eps2(j,1,c-3)), real(spinf*eps2(j,2,c-3))
			end if
		end do
		close(300+c)
	end do
	close(200)
	deallocate(ebse, a1, a2, eps1, eps2)
	deallocate(rvec,hopmatrices,ihopmatrices,ffactor)
end subroutine wtb_dielrawpol_files
