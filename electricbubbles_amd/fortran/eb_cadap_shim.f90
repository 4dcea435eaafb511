! eb_cadap_shim.f90 -- ISO_C_BINDING shim with the reference's external names.
!
! CADAPDR (cadapdrop.f:3-5), CADAPMOD (cadapmod.f:3-5) and CADAP (cadap.f:3-5) keep their
! 26-argument Fortran interface here and forward every argument, by reference, to the C-ABI
! of libeb.so (include/eb_cadap.h).  Linking this shim in place of cadapdrop.o, cadapmod.o and
! cadap.o lets the reference drivers (FASMVP, stokesbubb.f:569-571; INTERFACENEW,
! electric.f:835-837) run their FMM on the GPU without a source change.
!
! ND(K0:*) and H(K0:*) are passed on as the address of their element K0, as the reference's
! own callers do; COMPLEX*16 arrays travel as interleaved (re, im) doubles.
!
! DAPIF2 (dapif1.f:458-459) keeps its 18-argument interface the same way and forwards to
! eb_dapif2: FASMVP_LAP (laplace.f:214-215) and PVINTEV (pv_integral.f:181-182) then form their
! Hilbert-matrix products on the GPU.  Their CLOSE argument is an undeclared (REAL) variable;
! it is passed on by address and never called.

module eb_cadap_c
  use iso_c_binding, only: c_int32_t, c_double, c_double_complex
  implicit none
  interface
    integer(c_int32_t) function eb_cadapdr(iflag1, iflag2, levmax, xat, yat, dzdt, k0, kbod, nd, &
        dens, h, iwork, leniw, cwork, lenw, xtarg, ytarg, ntarg, gradw, phip, phipp, tol, inform, &
        ier, xleft, slength) bind(C, name="eb_cadapdr")
      import :: c_int32_t, c_double, c_double_complex
      integer(c_int32_t) :: iflag1, iflag2, levmax, k0, kbod, nd(*), iwork(*), leniw, lenw, ntarg
      integer(c_int32_t) :: inform(*), ier(*)
      real(c_double) :: xat(*), yat(*), h(*), xtarg(*), ytarg(*), tol, xleft, slength
      complex(c_double_complex) :: dzdt(*), dens(*), cwork(*), gradw(*), phip(*), phipp(*)
    end function
    integer(c_int32_t) function eb_cadapmod(iflag1, iflag2, levmax, xat, yat, dzdt, k0, kbod, nd, &
        dens, h, iwork, leniw, cwork, lenw, xtarg, ytarg, ntarg, gradw, phip, phipp, tol, inform, &
        ier, xleft, slength) bind(C, name="eb_cadapmod")
      import :: c_int32_t, c_double, c_double_complex
      integer(c_int32_t) :: iflag1, iflag2, levmax, k0, kbod, nd(*), iwork(*), leniw, lenw, ntarg
      integer(c_int32_t) :: inform(*), ier(*)
      real(c_double) :: xat(*), yat(*), h(*), xtarg(*), ytarg(*), tol, xleft, slength
      complex(c_double_complex) :: dzdt(*), dens(*), cwork(*), gradw(*), phip(*), phipp(*)
    end function
    integer(c_int32_t) function eb_cadap(iflag1, iflag2, levmax, xat, yat, dzdt, k0, kbod, nd, &
        dens, h, iwork, leniw, cwork, lenw, xtarg, ytarg, ntarg, gradw, phip, phipp, tol, inform, &
        ier, xleft, slength) bind(C, name="eb_cadap")
      import :: c_int32_t, c_double, c_double_complex
      integer(c_int32_t) :: iflag1, iflag2, levmax, k0, kbod, nd(*), iwork(*), leniw, lenw, ntarg
      integer(c_int32_t) :: inform(*), ier(*)
      real(c_double) :: xat(*), yat(*), h(*), xtarg(*), ytarg(*), tol, xleft, slength
      complex(c_double_complex) :: dzdt(*), dens(*), cwork(*), gradw(*), phip(*), phipp(*)
    end function
    integer(c_int32_t) function eb_dapif2(iout, iflag7, nat, napb, ninire, mex, ierr, inform, tol, &
        eps7, xa, ya, qa, poten, field, wksp, nsp, clos) bind(C, name="eb_dapif2")
      import :: c_int32_t, c_double, c_double_complex
      integer(c_int32_t) :: iout(*), iflag7, nat, napb, ninire, mex, ierr(*), inform(*), nsp
      real(c_double) :: tol, eps7, xa(*), ya(*), poten(*), wksp(*)
      complex(c_double_complex) :: qa(*), field(*)
      real :: clos
    end function
  end interface
end module eb_cadap_c

subroutine cadapdr(iflag1, iflag2, levmax, xat, yat, dzdt, k0, kbod, nd, dens, h, iwork, leniw, &
                   cwork, lenw, xtarg, ytarg, ntarg, gradw, phip, phipp, tol, inform, ier, xleft, &
                   slength)
  use eb_cadap_c
  implicit none
  integer(c_int32_t) :: iflag1, iflag2, levmax, k0, kbod, nd(*), iwork(*), leniw, lenw, ntarg
  integer(c_int32_t) :: inform(*), ier(*)
  real(c_double) :: xat(*), yat(*), h(*), xtarg(*), ytarg(*), tol, xleft, slength
  complex(c_double_complex) :: dzdt(*), dens(*), cwork(*), gradw(*), phip(*), phipp(*)
  integer(c_int32_t) :: rc
  rc = eb_cadapdr(iflag1, iflag2, levmax, xat, yat, dzdt, k0, kbod, nd, dens, h, iwork, leniw, &
                  cwork, lenw, xtarg, ytarg, ntarg, gradw, phip, phipp, tol, inform, ier, xleft, &
                  slength)
end subroutine cadapdr

subroutine cadapmod(iflag1, iflag2, levmax, xat, yat, dzdt, k0, kbod, nd, dens, h, iwork, leniw, &
                    cwork, lenw, xtarg, ytarg, ntarg, gradw, phip, phipp, tol, inform, ier, xleft, &
                    slength)
  use eb_cadap_c
  implicit none
  integer(c_int32_t) :: iflag1, iflag2, levmax, k0, kbod, nd(*), iwork(*), leniw, lenw, ntarg
  integer(c_int32_t) :: inform(*), ier(*)
  real(c_double) :: xat(*), yat(*), h(*), xtarg(*), ytarg(*), tol, xleft, slength
  complex(c_double_complex) :: dzdt(*), dens(*), cwork(*), gradw(*), phip(*), phipp(*)
  integer(c_int32_t) :: rc
  rc = eb_cadapmod(iflag1, iflag2, levmax, xat, yat, dzdt, k0, kbod, nd, dens, h, iwork, leniw, &
                   cwork, lenw, xtarg, ytarg, ntarg, gradw, phip, phipp, tol, inform, ier, xleft, &
                   slength)
end subroutine cadapmod

subroutine cadap(iflag1, iflag2, levmax, xat, yat, dzdt, k0, kbod, nd, dens, h, iwork, leniw, &
                 cwork, lenw, xtarg, ytarg, ntarg, gradw, phip, phipp, tol, inform, ier, xleft, &
                 slength)
  use eb_cadap_c
  implicit none
  integer(c_int32_t) :: iflag1, iflag2, levmax, k0, kbod, nd(*), iwork(*), leniw, lenw, ntarg
  integer(c_int32_t) :: inform(*), ier(*)
  real(c_double) :: xat(*), yat(*), h(*), xtarg(*), ytarg(*), tol, xleft, slength
  complex(c_double_complex) :: dzdt(*), dens(*), cwork(*), gradw(*), phip(*), phipp(*)
  integer(c_int32_t) :: rc
  rc = eb_cadap(iflag1, iflag2, levmax, xat, yat, dzdt, k0, kbod, nd, dens, h, iwork, leniw, &
                cwork, lenw, xtarg, ytarg, ntarg, gradw, phip, phipp, tol, inform, ier, xleft, &
                slength)
end subroutine cadap

subroutine dapif2(iout, iflag7, nat, napb, ninire, mex, ierr, inform, tol, eps7, xa, ya, qa, poten, &
                  field, wksp, nsp, clos)
  use eb_cadap_c
  implicit none
  integer(c_int32_t) :: iout(*), iflag7, nat, napb, ninire, mex, ierr(*), inform(*), nsp
  real(c_double) :: tol, eps7, xa(*), ya(*), poten(*), wksp(*)
  complex(c_double_complex) :: qa(*), field(*)
  real :: clos
  integer(c_int32_t) :: rc
  rc = eb_dapif2(iout, iflag7, nat, napb, ninire, mex, ierr, inform, tol, eps7, xa, ya, qa, poten, &
                 field, wksp, nsp, clos)
end subroutine dapif2
