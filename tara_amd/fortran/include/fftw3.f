! Include file for the reference's main programs when they are linked against
! tara_amd (libtara_dfftw.so in place of FFTW, libtara_gnucompat.so for the
! GNU runtime extensions they call).  The mains read it with
! include "fftw3.f" inside their declarations (e.g. src/main2csh.f95:11), so
! put this directory first on the include path.
!
! FFTW_ESTIMATE: the only planner flag the mains pass (the value of FFTW's
! Fortran interface; the shim accepts and ignores flags).
      INTEGER FFTW_ESTIMATE
      PARAMETER (FFTW_ESTIMATE=64)
! GNU Fortran's RAND extension, an intrinsic under gfortran but unknown to
! flang, so src/main2csh.f95:94-95 (implicit none) needs its type declared:
! a default-real function, provided by libtara_gnucompat.so.
      REAL, EXTERNAL :: rand
