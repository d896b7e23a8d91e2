# SavHip.jl — Julia binding of libsav_hip.so, the MI355X implementation of the SAV time
# step of Forstylime/Julia-Vesicles-growth-and-shrinkage.
#
# Include it after src/Types.jl (it uses Config and FieldState) and call
# SavHip.run_simulation with the arguments of src/simulation.jl:2-4.  The library
# replaces the loop body (solve_step1..7, update_state!, the back-transforms and the
# monitor, src/simulation.jl:73-110,146-148); the schedule, the snapshot rule and the
# returned histories stay the reference's.  Arrays cross the boundary as they are in
# Julia (column-major, x fastest; ComplexF64 is interleaved double[2]); the library
# copies during each call, so GC.@preserve is all the lifetime management needed.
#
# No Julia toolchain exists where this file was written; the same C ABI is exercised
# from Python (sav.py, ctypes) and from C++ (abi_client.cpp) by the tests.
module SavHip

using Printf

const LIB = get(ENV, "SAV_HIP_LIB", joinpath(@__DIR__, "..", "libsav_hip.so"))

const SAV_PRESENT = Cint(0)
const SAV_OLD = Cint(1)
const SAV_W_KRYLOV_NOT_CONVERGED = Cint(1)
const SAV_E_WC_NEGATIVE = Cint(-3)
const SAV_PREC_FP64 = Cint(0)      # the reference's arithmetic
const SAV_PREC_FP32 = Cint(1)      # FP32 fields and transforms, FP64 reductions
const SAV_PREC_MIXED = Cint(2)     # FP64 except the psi solve's inner iterations (FP32, FP64 defect correction)

struct SavConfig                      # sav_config_t, field for field
    ndim::Int32; N::Int32; Nx::Int32; Ny::Int32; Nz::Int32
    Lx::Float64; Ly::Float64; Lz::Float64
    epsilon::Float64; M_phi::Float64; M0_psi::Float64; eta::Float64; gamma_surf::Float64
    gamma_area::Int64; gamma_bend::Float64
    gamma_in::Int64; beta_in::Int64; gamma_out::Int64; beta_out::Int64; lamda::Int64
    psi_in_v::Ptr{Float64}; psi_out_v::Ptr{Float64}
    S1::Float64; S2::Float64; S3::Float64; S4::Float64; C1::Float64; C2::Float64; C3::Float64
    dt::Float64; T::Float64; tol::Float64
    precision::Int32; precond::Int32; itmax::Int32; atol::Float64; device::Int32
end

struct StepStats                      # sav_step_stats_t
    krylov_iters::NTuple{3,Int32}; krylov_solved::NTuple{3,Int32}; krylov_rnorm::NTuple{3,Float64}
    energy::Float64; area_ratio::Float64; Q::Float64; R1::Float64; R2::Float64; R3::Float64; status::Int32
end

last_error(h) = unsafe_string(ccall((:sav_last_error, LIB), Cstring, (Ptr{Cvoid},), h))

"""Device handle for `conf` (Config of src/Types.jl:27-74): sav_create.
precision SAV_PREC_FP64 (the reference's arithmetic), SAV_PREC_MIXED (FP64 except the psi
solve's inner iterations, same stopping test) or SAV_PREC_FP32 (FP32 fields, FP64 reductions)."""
function create(conf; precision::Integer=0, itmax::Integer=20, atol::Real=1e-14, device::Integer=0)
    pin = collect(Float64, conf.psi_in_v)
    pout = collect(Float64, conf.psi_out_v)
    ndim = hasproperty(conf, :Nz) ? 3 : 2
    GC.@preserve pin pout begin
        c = SavConfig(ndim, conf.N, conf.Nx, conf.Ny, ndim == 3 ? conf.Nz : 1,
                      conf.Lx, conf.Ly, ndim == 3 ? conf.Lz : 1.0,
                      conf.epsilon, conf.M_phi, conf.M0_psi, conf.eta, conf.gamma_surf,
                      conf.gamma_area, conf.gamma_bend, conf.gamma_in, conf.beta_in,
                      conf.gamma_out, conf.beta_out, conf.lamda, pointer(pin), pointer(pout),
                      conf.S1, conf.S2, conf.S3, conf.S4, conf.C1, conf.C2, conf.C3,
                      conf.dt, conf.T, conf.tol, precision, 0, itmax, atol, device)
        h = Ref{Ptr{Cvoid}}(C_NULL)
        rc = ccall((:sav_create, LIB), Cint, (Ref{SavConfig}, Ref{Ptr{Cvoid}}), c, h)
        rc == 0 || error("sav_create: ", last_error(C_NULL))
        return h[]
    end
end

destroy(h) = ccall((:sav_destroy, LIB), Cint, (Ptr{Cvoid},), h)

"""Vesicle count N of a FieldState: the last dimension of phi, (Nx, Ny, N) or (Nx, Ny, Nz, N)."""
nvesicles(s) = size(s.phi, ndims(s.phi))

"""Copies a FieldState (src/Types.jl:167-201) into the handle (which = SAV_PRESENT/SAV_OLD)."""
function set_state!(h, which, s)
    sc = Float64[s.Q, s.R1, s.R2, s.R3]
    # the library reads N values of A0; the reference's FieldState holds A0 = [0.0]
    # (Init.jl:215, Types.jl:228) until the SAV initialisation, so a short vector is padded
    N = nvesicles(s)
    A0 = zeros(N)
    copyto!(A0, 1, s.A0, 1, min(N, length(s.A0)))
    GC.@preserve s sc A0 begin
        rc = ccall((:sav_set_state, LIB), Cint,
                   (Ptr{Cvoid}, Cint, Ptr{Float64}, Ptr{ComplexF64}, Ptr{Float64}, Ptr{ComplexF64},
                    Ptr{Float64}, Ptr{ComplexF64}, Ptr{Float64}, Ptr{ComplexF64},
                    Ptr{Float64}, Ptr{ComplexF64}, Ptr{Float64}, Ptr{ComplexF64},
                    Ptr{Float64}, Ptr{Float64}, Float64),
                   h, which, s.phi, s.phi_hat, s.psi, s.psi_hat, s.u, s.u_hat, s.p, s.p_hat,
                   s.mu, s.mu_hat, s.nu, s.nu_hat, sc, A0, s.dt)
        rc == 0 || error("sav_set_state: ", last_error(h))
    end
    return s
end

"""Copies the handle's state into `s` (same shapes as FieldState(Nx, Ny, N))."""
function get_state!(h, which, s)
    sc = zeros(4)
    A0 = zeros(nvesicles(s))  # the library writes N values
    dt = Ref{Float64}(0.0)
    GC.@preserve s sc A0 begin
        rc = ccall((:sav_get_state, LIB), Cint,
                   (Ptr{Cvoid}, Cint, Ptr{Float64}, Ptr{ComplexF64}, Ptr{Float64}, Ptr{ComplexF64},
                    Ptr{Float64}, Ptr{ComplexF64}, Ptr{Float64}, Ptr{ComplexF64},
                    Ptr{Float64}, Ptr{ComplexF64}, Ptr{Float64}, Ptr{ComplexF64},
                    Ptr{Float64}, Ptr{Float64}, Ref{Float64}),
                   h, which, s.phi, s.phi_hat, s.psi, s.psi_hat, s.u, s.u_hat, s.p, s.p_hat,
                   s.mu, s.mu_hat, s.nu, s.nu_hat, sc, A0, dt)
        rc == 0 || error("sav_get_state: ", last_error(h))
    end
    s.Q, s.R1, s.R2, s.R3 = sc
    which == SAV_PRESENT && (s.A0 = A0)
    return s
end

"""SAV initialisation of src/simulation.jl:24-53 on the present state; old = present."""
function init_sav!(h)
    rc = ccall((:sav_init_sav, LIB), Cint, (Ptr{Cvoid},), h)
    rc == SAV_E_WC_NEGATIVE && error("FATAL: W + C < 0")
    rc == 0 || error("sav_init_sav: ", last_error(h))
end

"""One loop body (src/simulation.jl:73-110) with loop index n; returns the StepStats."""
function step!(h, dt, n)
    st = Ref{StepStats}()
    rc = ccall((:sav_step, LIB), Cint, (Ptr{Cvoid}, Float64, Cint, Ref{StepStats}), h, dt, n, st)
    rc == SAV_E_WC_NEGATIVE && error("FATAL: W + C < 0")
    rc == SAV_W_KRYLOV_NOT_CONVERGED && @warn "BiCGSTAB did not converge"
    rc < 0 && error("sav_step: ", last_error(h))
    return st[]
end

"""run_simulation of src/simulation.jl:2-155 with the loop body on the GPU.  Needs the
reference's set_para_base, build_operators, generate_initial_condition, FieldState
and MAT.matwrite in scope (include this file after them)."""
function run_simulation(dt_val::Float64, T_val::Float64, state_type::Int;
                        save_path::String="./results", save_interval::Float64=0.001,
                        precision::Integer=0)
    mat_folder = joinpath(save_path, "MAT", "case3")
    mkpath(mat_folder)
    conf = Main.set_para_base(dt_val, T_val)
    ops = Main.build_operators(conf)
    t1 = range(0.0, 1e-5, step=1e-7)
    t2 = range(1e-5, T_val, step=conf.dt)
    Dt = vcat(t1, t2[2:end])
    Nt = length(Dt)
    present = Main.generate_initial_condition(conf, ops, state_type)
    h = create(conf; precision=precision)
    try
        set_state!(h, SAV_PRESENT, present)
        init_sav!(h)
        energy_history = Float64[]
        area_ratio_history = Float64[]
        last_save_idx = -1
        for n in 1:Nt-1
            st = step!(h, Dt[n+1] - Dt[n], n)
            curr_save_idx = floor(Int, (Dt[n] + 1e-12) / save_interval)
            if curr_save_idx > last_save_idx
                get_state!(h, SAV_PRESENT, present)
                t_str = @sprintf("%.2e", Dt[n])
                d = Dict("phi" => present.phi, "psi" => present.psi, "mu" => present.mu,
                         "nu" => present.nu, "p" => present.p, "u" => present.u,
                         "R1" => present.R1, "R2" => present.R2, "R3" => present.R3, "Q" => present.Q)
                Main.matwrite(joinpath(mat_folder, "state_$(ndims(present.phi) - 1)d_t$(t_str).mat"), d)
                last_save_idx = curr_save_idx
            end
            push!(energy_history, st.energy)
            push!(area_ratio_history, st.area_ratio)
        end
        get_state!(h, SAV_PRESENT, present)
        return present, energy_history, area_ratio_history, Dt
    finally
        destroy(h)
    end
end

end # module
