! Test infrastructure only (part of the oracle): calls the reference's own k-grid
! density routines rkmesh and rkmesh2D (subroutines/mhkpack_subs.f90:1-52), compiled
! from the reference sources where they lie (oracle/Makefile), on cases read from
! stdin, one per line: rk followed by the nine entries of rlat(3,3) row by row
! (rows = lattice vectors a1, a2, a3).  Prints, per case, the RK3D grid then the
! RK2D grid.  Nothing in the product loads or runs this program.

INCLUDE "subroutines/general_subs.f90"
INCLUDE "subroutines/mhkpack_subs.f90"

program ref_rkmesh
	implicit none
	double precision :: rk
	double precision, dimension(3,3) :: rlat
	integer, dimension(3) :: g3, g2
	integer :: ios
	do
		read(*,*,iostat=ios) rk, rlat(1,1), rlat(1,2), rlat(1,3), rlat(2,1), rlat(2,2), rlat(2,3), &
			rlat(3,1), rlat(3,2), rlat(3,3)
		if (ios /= 0) exit
		call rkmesh(rk, rlat, g3)
		call rkmesh2D(rk, rlat, g2)
		write(*,'(6I8)') g3, g2
	end do
end program ref_rkmesh
