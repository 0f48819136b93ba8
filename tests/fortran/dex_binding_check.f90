module dexmod
  use iso_c_binding
  implicit none
  type, bind(C) :: dex_params
     integer(c_int32_t) :: n
     real(c_double)     :: delta, theta
     integer(c_int32_t) :: charge, nsmall, tfactor, precision, device
  end type
  type, bind(C) :: dex_solve_opts
     integer(c_int32_t) :: strict_cholesky, verbose, refresh_every, fp64_restarts
     real(c_double)     :: fp32_tol
     integer(c_int32_t) :: symmetry, lock, inner_block, guard, inner_fp32
  end type
  type, bind(C) :: dex_result
     integer(c_int32_t) :: iterations, converged
     integer(c_int64_t) :: matvecs
     real(c_double)     :: max_residual, t_total, t_matvec, t_prec, t_gram, t_update, t_smalleig
     integer(c_int32_t) :: fallbacks, fp64_restarts_used
  end type
  interface
     subroutine dex_params_default(p) bind(C, name="dex_params_default")
       import; type(dex_params) :: p
     end subroutine
     subroutine dex_solve_opts_default(o) bind(C, name="dex_solve_opts_default")
       import; type(dex_solve_opts) :: o
     end subroutine
     integer(c_int) function dex_ctx_create(p, ctx) bind(C, name="dex_ctx_create")
       import; type(dex_params), intent(in) :: p; type(c_ptr) :: ctx
     end function
     integer(c_int) function dex_ctx_destroy(ctx) bind(C, name="dex_ctx_destroy")
       import; type(c_ptr), value :: ctx
     end function
     ! MYMATMUL(IN,OUT,HOWMANY,CHECKSIZE)
     integer(c_int) function dex_matvec(ctx, in, out, howmany, checksize) bind(C, name="dex_matvec")
       import; type(c_ptr), value :: ctx
       complex(c_double_complex), intent(in)  :: in(*)
       complex(c_double_complex), intent(out) :: out(*)
       integer(c_int32_t), value :: howmany
       integer(c_int64_t), value :: checksize
     end function
     ! MYDOTSUB(BRAS,KETS,SIZE,NN,MM,OUTDOTS)
     integer(c_int) function dex_dots(ctx, bras, kets, size, nn, mm, out) bind(C, name="dex_dots")
       import; type(c_ptr), value :: ctx
       complex(c_double_complex), intent(in) :: bras(*), kets(*)
       integer(c_int64_t), value :: size
       integer(c_int32_t), value :: nn, mm
       complex(c_double_complex), intent(out) :: out(*)
     end function
     ! GPLMR_EIG(HAM,PARAMS,MPIDATA,GMSOLVER,X,E,M,MB,SIGMAIN,NEV,NS,MAXIT,TOL,PRECONDITIONER,OUTMODULUS)
     integer(c_int) function dex_gplmr_eig(ctx, x, e, m, mb, sre, sim, nev, ns, maxit, tol, inner, &
          outmodulus, opts, res) bind(C, name="dex_gplmr_eig")
       import; type(c_ptr), value :: ctx
       complex(c_double_complex), intent(inout) :: x(*)
       complex(c_double_complex), intent(out)   :: e(*)
       integer(c_int32_t), value :: m, mb, nev, ns, maxit, inner, outmodulus
       real(c_double), value :: sre, sim, tol
       type(dex_solve_opts), intent(in) :: opts
       type(dex_result), intent(out) :: res
     end function
  end interface
end module dexmod

module dexhostmod
  use iso_c_binding
  implicit none
  interface
     integer(c_int) function dex_host_kematrix(nn, delta, t) bind(C, name="dex_host_kematrix")
       import; integer(c_int32_t), value :: nn; real(c_double), value :: delta; real(c_double) :: t(*)
     end function
     integer(c_int) function dex_host_getinverse(nout, delta, nsmall, tfactor, tinv) bind(C, name="dex_host_getinverse")
       import; integer(c_int32_t), value :: nout, nsmall, tfactor; real(c_double), value :: delta
       real(c_double) :: tinv(*)
     end function
  end interface
end module dexhostmod

program dex_binding_check
  use dexmod
  use dexhostmod
  implicit none
  type(dex_params) :: p
  type(dex_solve_opts) :: o
  real(c_double) :: t(5,5), tinv(-5:5,-5:5,-5:5)
  integer :: st
  call dex_params_default(p)
  call dex_solve_opts_default(o)
  st = dex_host_kematrix(5, 0.5d0, t)
  if (st /= 0) stop 1
  st = dex_host_getinverse(5, 0.5d0, 22, 4, tinv)
  if (st /= 0) stop 2
  write(*,'(I0,1X,I0,1X,I0,1X,F20.16,1X,F20.16,1X,F20.16)') p%n, p%nsmall, o%fp64_restarts, t(1,1), t(2,1), tinv(3,0,0)
end program dex_binding_check
