! range_table.f90 -- prints the block distribution of the drop-in's range
! (GHOST fftp-3/fftp.f90:17-48) over a sweep of its arguments, one line
! "n1 n2 nprocs irank ista iend" per call, for the CPU test of range.
PROGRAM range_table
  IMPLICIT NONE
  INTEGER :: n1, n2, np, r, ista, iend, q
  INTEGER, PARAMETER :: starts(3) = (/1, 0, -3/)
  INTEGER, PARAMETER :: slabs(4) = (/257, 512, 256, 129/)
  DO q = 1, 3
    n1 = starts(q)
    DO n2 = n1 - 1, n1 + 40
      DO np = 1, 12
        DO r = 0, np - 1
          CALL range(n1, n2, np, r, ista, iend)
          WRITE(*, '(6I8)') n1, n2, np, r, ista, iend
        END DO
      END DO
    END DO
  END DO
  ! GHOST's own slab sizes: kx columns nx/2+1 and z planes nz of the
  ! 512^3 and 512x512x256 grids over 1..64 ranks
  DO n2 = 1, 4
    DO np = 1, 64
      DO r = 0, np - 1
        CALL range(1, slabs(n2), np, r, ista, iend)
        WRITE(*, '(6I8)') 1, slabs(n2), np, r, ista, iend
      END DO
    END DO
  END DO
END PROGRAM range_table
