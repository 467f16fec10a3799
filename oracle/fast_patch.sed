# Text patch of the "fast" reference build (applied by oracle/Makefile to a
# temporary copy of SFIT_DVR3D/H2S/dvr3drjz.f90 outside the repository).
# 1. mkham3 loops 83/82/84 are multiplied by rm2t, which is 0 for the
#    theta-first zquad2 runs (dvr3drjz.f90:2390, 2424-2438): guarded so that
#    they run only for rm2t /= 0.  Skipped, they would only have added +-0.
/^ *do 83 na = 1,npnta/i\               if (rm2t .ne. 0.0d0) then
/^ *83 *continue/a\               endif
# 2. dsyev with its optimal workspace (oracle/dsyev_optlw.f90)
s/call dsyev *(/call dsyev_optlw(/
