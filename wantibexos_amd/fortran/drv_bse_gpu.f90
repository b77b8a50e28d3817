! The BSE + spectrum and exciton-band paths of the reference's program main
! (wtb_main.f90:60-190, 253-326) with the GPU drop-ins of bse_solver_gpu.f90 in place of the
! reference's bsesolver, bsedielraw and spdielraw (and their variants), and the drop-ins of
! optprop_gpu.f90 in place of the post-processing of the spectra and the exciton lifetimes
! (bseoptprop, spoptprop, lifetime and their -pol forms).  The input parser and the TB
! reader are the reference's own, INCLUDEd from its sources by the Makefile.
!
!   wtb_gpu.x < input.dat     (the reference's KEY= value input)

INCLUDE "subroutines/general_subs.f90"
INCLUDE "subroutines/mhkpack_subs.f90"
INCLUDE "subroutines/module_input_read.f90"
INCLUDE "subroutines/coulomb_pot.f90"
INCLUDE "subroutines/bse_subs.f90"
INCLUDE "wtb_gpu.f90"
INCLUDE "bse_solver_gpu.f90"
INCLUDE "optprop_gpu.f90"

program wtb_gpu_main
	use input_variables
	use hamiltonian_input_variables
	implicit none
	integer :: erro
	double precision,dimension(3,3) :: rlatv

	call input_read

	OPEN(UNIT=2055, FILE= params,STATUS='old', IOSTAT=erro)
	if (erro/=0) stop "Error opening wannier90 hr input file"
	read(2055,*) systype
	read(2055,*) scs
	read(2055,*) efermi
	read(2055,*) rlatv(1,1),rlatv(1,2),rlatv(1,3)
	read(2055,*) rlatv(2,1),rlatv(2,2),rlatv(2,3)
	read(2055,*) rlatv(3,1),rlatv(3,2),rlatv(3,3)
	if (meshtype .eq. "RK3D") then
		call rkmesh(rk,rlatv,ngrid)
	else if (meshtype .eq. "RK2D") then
		call rkmesh2D(rk,rlatv,ngrid)
	end if
	close(2055)

	if (bse .and. bset) then
		write(*,*) "ERROR: Choose BSE=T or BSET=T not both"
		stop
	end if
	if (bse) then
		call bsesolver(nthreads,outputfolder,calcparms,ngrid,nc,nv,numdos, &
			ebse0,ebsef,numbse,cshift,ktol,params,kpaths,kpathsbse,orbw,ediel, &
			exc,mshift,coultype,ez,w,r0,lc,rk,meshtype,bsewf,excwf0,excwff,dtfull,cpol,tmcoef)
	else if (bset) then
		call bsesolvertemp(nthreads,outputfolder,calcparms,ngrid,nc,nv,numdos, &
			ebse0,ebsef,numbse,cshift,ktol,params,kpaths,kpathsbse,orbw,ediel, &
			exc,mshift,coultype,ez,w,r0,lc,rk,meshtype,bsewf,excwf0,excwff,dtfull,&
			cpol,tmcoef,st,phavg,ta,temp)
	end if
	if (bsekpath .and. bsetbnd) then
		write(*,*) "ERROR: Choose BSE_BND=T or BSET_BND=T not both"
		stop
	else if (bsekpath) then
		call bsebnds(nthreads,outputfolder,calcparms,ngrid,nc,nv,numdos, &
			ebse0,ebsef,numbse,sme,ktol,params,kpaths,kpathsbse,orbw,ediel, &
			exc,mshift,coultype,ez,w,r0,lc,rk,meshtype,bsewf,excwf0,excwff)
	else if (bsetbnd) then
		call bsebndstemp(nthreads,outputfolder,calcparms,ngrid,nc,nv,numdos, &
			ebse0,ebsef,numbse,sme,ktol,params,kpaths,kpathsbse,orbw,ediel, &
			exc,mshift,coultype,ez,w,r0,lc,rk,meshtype,bsewf,excwf0,excwff,&
			st,phavg,ta,temp)
	end if
	! after the exciton bands, as in wtb_main.f90:209-218 (with BSET= T it replaces the
	! finite-temperature sp_optics.dat by the T = 0 one, as there)
	if (spdiel) then
		call spoptics(nthreads,outputfolder,calcparms,ngrid,nc,nv,numdos, &
			ebse0,ebsef,numbse,cshift,ktol,params,kpaths,kpathsbse,orbw,ediel, &
			exc,mshift,coultype,rk,meshtype,tmcoef)
	end if
	! the spectrum block of wtb_main.f90:257-326
	if (spec .and. (bse .or. bset)) then
		call bsedielraw(nthreads,outputfolder,renorm,params,ngrid,nc,nv,ebse0,ebsef,numbse,cshift)
		call bseoptprop(numbse,outputfolder)
		if (.not. bset) call lifetime(sysdim,numbse,ngrid,rlat,nc,nv,outputfolder)
		if (cpol) then
			call bsedielrawpol(nthreads,outputfolder,renorm,params,ngrid,nc,nv,ebse0,ebsef, &
					   numbse,cshift)
			call bseoptproppol(numbse,outputfolder)
			if (.not. bset) call lifetimepol(sysdim,numbse,ngrid,rlat,nc,nv,outputfolder)
		end if
	end if
	if (spec .and. (spdiel .or. bse .or. bset)) then
		call spdielraw(nthreads,outputfolder,renorm,params,ngrid,nc,nv,ebse0,ebsef,numbse,cshift)
		call spoptprop(numbse,outputfolder)
	end if
	if (spec .and. (cpol .or. bse .or. bset)) then
		call spdielrawpol(nthreads,outputfolder,renorm,params,ngrid,nc,nv,ebse0,ebsef,numbse,cshift)
		call spoptproppol(numbse,outputfolder)
	end if
end program wtb_gpu_main
