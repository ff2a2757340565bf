! Test infrastructure only (a checker, never the product): prints numbers
! with the Fortran edit descriptors of the KPP-Standalone state file
! (GeosCore/kppsa_interface_mod.F90:750-838), formatted by the Fortran
! runtime, so the library's own formatting can be compared with it.
! Input on stdin, one per line: "R <int64 IEEE-754 bit pattern>" for a real
! or "I <integer>" for an integer.  Output: one line per descriptor.
PROGRAM kppsa_fmt
  IMPLICIT NONE
  CHARACTER(LEN=1) :: tag
  INTEGER(8)       :: bits
  REAL(8)          :: x
  INTEGER          :: ios
  DO
     READ( *, *, IOSTAT=ios ) tag, bits
     IF ( ios /= 0 ) EXIT
     IF ( tag == 'R' ) THEN
        x = TRANSFER( bits, x )
        WRITE( *, '(es25.16e3)' ) x
        WRITE( *, '(es10.2e2)'  ) x
        WRITE( *, '(f11.4)'     ) x
        WRITE( *, '(f11.2)'     ) x
        WRITE( *, '(e11.4)'     ) x
        WRITE( *, '(F13.6)'     ) x
     ELSE
        WRITE( *, '(i6)' ) INT( bits )
     ENDIF
  ENDDO
END PROGRAM kppsa_fmt
