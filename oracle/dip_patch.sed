# Text patch of the reference DIPOLE3 build (applied by oracle/Makefile to a
# temporary copy of source/dipole3_segmented.f90 outside the repository).
# 1. The Intel-only recordtype='segmented' keyword is dropped from the OPEN
#    statements (amdflang rejects it); the files are then default sequential
#    unformatted, the format every other program of this build writes.
s/, *recordtype *= *'segmented'//
# 2. The file ends with an example DIPD for AB2 molecules that calls a user
#    DIPSA; the dipole surface is a link-time user routine, so the example is
#    cut and the surface of each build is linked instead (the reference's own
#    H3+ test does the same with H3+Jacobi/source/dip_h3p.f).
/^      SUBROUTINE DIPD(DIPC,R1,R2,XCOS,NU)/,$d
