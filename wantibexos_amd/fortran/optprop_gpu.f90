! Drop-ins for the post-processing of the spectra and the exciton lifetimes, with the
! reference's names, argument lists, input files and output files:
!   bseoptprop, spoptprop, bseoptproppol, spoptproppol
!       (subprograms/diel-pp-bse.f90:3-173, diel-pp.f90, diel-pp-bse-pol.f90,
!        diel-pp-pol.f90): refractive index, extinction coefficient, reflectivity,
!        absorption coefficient and energy-loss function of every spectrum, on the GPU
!        through wtb_optprop;
!   lifetime, lifetimepol (subprograms/exciton_lifetime.f90:2-187): radiative lifetimes
!        of every exciton, on the GPU through wtb_exciton_lifetime.
! The files are read and written here exactly as the reference reads and writes them.

! the five optical-constant files of one spectrum set: ncomp input files (names in the
! reference's reading order), output columns in the order perm(1:ncomp) of the inputs
subroutine wtb_optprop_files(numbse, outputfolder, ncomp, innames, perm, outpref, outsuf, header)
	use wtb_gpu
	implicit none
	double precision :: numbse
	character(len=70) :: outputfolder
	integer :: ncomp
	character(len=*), dimension(ncomp) :: innames
	integer, dimension(ncomp) :: perm
	character(len=*) :: outpref, outsuf
	character(len=*), dimension(ncomp) :: header
	character(len=24), dimension(5) :: qname = [character(len=24) :: "refractive_index", "extinction_coef", &
		"reflectibility", "absorption_coef", "en_loss_func"]
	character(len=2) :: aread
	double precision :: flag
	double precision, allocatable :: energy(:), er(:,:), ei(:,:), out(:,:,:)
	integer :: dimpp, i, c, q, erro, u
	type(c_ptr) :: ctx
	integer(c_int) :: rc
	character(len=16) :: fmt

	dimpp = int(numbse)
	allocate(energy(dimpp), er(dimpp,ncomp), ei(dimpp,ncomp), out(dimpp,ncomp,5))
	do c = 1, ncomp
		OPEN(UNIT=99+c, FILE=trim(outputfolder)//trim(innames(c)), STATUS='old', IOSTAT=erro)
		if (erro/=0) then
			write(*,*) "Error opening ", trim(innames(c)), " input file"
			stop
		end if
		read(99+c,*) aread
	end do
	do i = 1, dimpp
		read(100,*) energy(i), er(i,1), ei(i,1)
		do c = 2, ncomp
			read(99+c,*) flag, er(i,c), ei(i,c)
		end do
	end do
	do c = 1, ncomp
		close(99+c)
	end do

	rc = wtb_create(0_c_int, ctx)
	call wtb_check(rc)
	rc = wtb_optprop(ctx, int(dimpp,c_int), int(ncomp,c_int), energy, er, ei, out)
	call wtb_check(rc)
	rc = wtb_destroy(ctx)

	write(fmt, "(A1,I1,A6)") "(", ncomp + 1, "F15.4)"
	do q = 1, 5
		u = 100*(q + 1)
		OPEN(UNIT=u, FILE=trim(outputfolder)//trim(outpref)//trim(qname(q))//trim(outsuf)//".dat", &
		     STATUS='unknown', IOSTAT=erro)
		if (erro/=0) then
			write(*,*) "Error opening ", trim(outpref)//trim(qname(q))//trim(outsuf), " output file"
			stop
		end if
		if (ncomp == 6) then
			write(u,*) "#","  ","energy","  ",trim(header(1)),"  ",trim(header(2)),"  ",trim(header(3)),"  ", &
				trim(header(4)),"  ",trim(header(5)),"  ",trim(header(6))
		else
			write(u,*) "#","  ","energy","  ",trim(header(1)),"  ",trim(header(2)),"  ",trim(header(3)),"  ", &
				trim(header(4)),"  ",trim(header(5))
		end if
	end do
	do i = 1, dimpp
		do q = 1, 5
			write(100*(q + 1),fmt) energy(i), (out(i,perm(c),q), c = 1, ncomp)
		end do
	end do
	do q = 1, 5
		close(100*(q + 1))
	end do
	deallocate(energy, er, ei, out)
end subroutine wtb_optprop_files

subroutine bseoptprop(numbse,outputfolder)
	implicit none
	character(len=70) :: outputfolder
	double precision :: numbse
	call wtb_optprop_files(numbse, outputfolder, 6, [character(len=20) :: "bse_diel_xx.dat", "bse_diel_xy.dat", &
		"bse_diel_xz.dat", "bse_diel_yy.dat", "bse_diel_yz.dat", "bse_diel_zz.dat"], [1, 4, 6, 2, 3, 5], &
		"bse_", "", [character(len=2) :: "xx", "yy", "zz", "xy", "xz", "yz"])
end subroutine bseoptprop

subroutine spoptprop(numbse,outputfolder)
	implicit none
	character(len=70) :: outputfolder
	double precision :: numbse
	call wtb_optprop_files(numbse, outputfolder, 6, [character(len=20) :: "sp_diel_xx.dat", "sp_diel_xy.dat", &
		"sp_diel_xz.dat", "sp_diel_yy.dat", "sp_diel_yz.dat", "sp_diel_zz.dat"], [1, 4, 6, 2, 3, 5], &
		"sp_", "", [character(len=2) :: "xx", "yy", "zz", "xy", "xz", "yz"])
end subroutine spoptprop

subroutine bseoptproppol(numbse,outputfolder)
	implicit none
	character(len=70) :: outputfolder
	double precision :: numbse
	call wtb_optprop_files(numbse, outputfolder, 5, [character(len=20) :: "bse_diel-pol_x.dat", &
		"bse_diel-pol_y.dat", "bse_diel-pol_z.dat", "bse_diel-pol_sp.dat", "bse_diel-pol_sm.dat"], &
		[1, 2, 3, 4, 5], "bse_", "-pol", [character(len=2) :: "x", "y", "z", "sp", "sm"])
end subroutine bseoptproppol

subroutine spoptproppol(numbse,outputfolder)
	implicit none
	character(len=70) :: outputfolder
	double precision :: numbse
	call wtb_optprop_files(numbse, outputfolder, 5, [character(len=20) :: "sp_diel-pol_x.dat", &
		"sp_diel-pol_y.dat", "sp_diel-pol_z.dat", "sp_diel-pol_sp.dat", "sp_diel-pol_sm.dat"], &
		[1, 2, 3, 4, 5], "sp_", "-pol", [character(len=2) :: "x", "y", "z", "sp", "sm"])
end subroutine spoptproppol

! lifetime and lifetimepol: read the exciton table (energy + ncomp activities, F15.4),
! lifetimes of the summed x+y+z strength and of nout components, written "(1F15.4,nE15.4)"
subroutine wtb_lifetime_files(sysdim,ngrid,rlat,nc,nv,outputfolder,inname,outname,ncomp,nout,header)
	use wtb_gpu
	implicit none
	character(len=5) :: sysdim
	integer, dimension(3) :: ngrid
	double precision, dimension(3,3) :: rlat
	integer :: nc, nv, ncomp, nout
	character(len=70) :: outputfolder
	character(len=*) :: inname, outname, header
	character(len=70) :: cflag
	double precision, allocatable :: fosc(:,:), lft(:,:)
	integer :: dimbse, i, c, erro, isd
	integer(c_int) :: gr(3)
	type(c_ptr) :: ctx
	integer(c_int) :: rc
	character(len=20) :: fmt

	OPEN(UNIT=100, FILE=trim(outputfolder)//trim(inname), STATUS='old', IOSTAT=erro)
	if (erro/=0) then
		write(*,*) "Error opening ", trim(inname), " input file"
		stop
	end if
	OPEN(UNIT=300, FILE=trim(outputfolder)//trim(outname), STATUS='unknown', IOSTAT=erro)
	if (erro/=0) then
		write(*,*) "Error opening ", trim(outname), " output file"
		stop
	end if
	dimbse = ngrid(1)*ngrid(2)*ngrid(3)*nc*nv
	allocate(fosc(dimbse,ncomp+1), lft(dimbse,nout+1))
	read(100,*) cflag
	do i = 1, dimbse
		read(100,*) (fosc(i,c), c = 1, ncomp + 1)
	end do
	write(300,*) header
	select case (sysdim)
	case ("3D"); isd = 3
	case ("2D"); isd = 2
	case ("1D"); isd = 1
	case default
		write(*,*) "Wrong System dimension"
		STOP
	end select
	gr = ngrid
	rc = wtb_create(0_c_int, ctx)
	call wtb_check(rc)
	rc = wtb_exciton_lifetime(ctx, int(isd,c_int), gr, rlat, int(dimbse,c_int), int(ncomp,c_int), &
				  int(nout,c_int), fosc, lft)
	call wtb_check(rc)
	rc = wtb_destroy(ctx)
	write(fmt, "(A9,I1,A6)") "(1F15.4,", nout + 1, "E15.4)"
	do i = 1, dimbse
		write(300,fmt) fosc(i,1), (lft(i,c), c = 1, nout + 1)
	end do
	deallocate(fosc, lft)
	close(100)
	close(300)
end subroutine wtb_lifetime_files

subroutine lifetime(sysdim,numbse,ngrid,rlat,nc,nv,outputfolder)
	implicit none
	integer :: nc,nv
	double precision :: numbse
	integer,dimension(3) :: ngrid
	double precision,dimension(3,3) :: rlat
	character(len=70) :: outputfolder
	character(len=5) :: sysdim
	call wtb_lifetime_files(sysdim,ngrid,rlat,nc,nv,outputfolder,"bse_opt_diel.dat","exciton_lifetime.dat", &
		6, 3, "#energy(eV) sum-lft(s)  x-lft(s) y-lft(s) z-lft(s)")
end subroutine lifetime

subroutine lifetimepol(sysdim,numbse,ngrid,rlat,nc,nv,outputfolder)
	implicit none
	integer :: nc,nv
	double precision :: numbse
	integer,dimension(3) :: ngrid
	double precision,dimension(3,3) :: rlat
	character(len=70) :: outputfolder
	character(len=5) :: sysdim
	call wtb_lifetime_files(sysdim,ngrid,rlat,nc,nv,outputfolder,"bse_opt_diel-pol.dat", &
		"exciton_lifetime-pol.dat", 5, 5, &
		"#energy(eV) sum-lft(s)  x-lft(s) y-lft(s) z-lft(s) sp-lft(s) sm-lft(s)")
end subroutine lifetimepol
