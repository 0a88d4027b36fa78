# hip.jl -- MI355X backend for ThreePhotons.jl, included by the module root in
# place of (or after) cuda.jl, exactly as cuda.jl is included today
# (src/ThreePhotons.jl:145-150).  Every method calls libtp.so through ccall;
# no CUDA/CUBLAS package is loaded.
#
#   HIP_init()                         replaces CUDA_init            (cuda.jl:10-21)
#   BasisTypeHip <: AbstractBasisType  replaces BasisTypeCuda        (cuda.jl:24-42)
#   HIP_calculate_basis(...)           replaces CUDA_calculate_basis (cuda.jl:57-69)
#   FullCorrelation_parallized(::SphericalHarmonicsVolume, ::BasisTypeHip, ...)   (cuda.jl:74-108)
#   energy(::SphericalHarmonicsVolume, ::BasisTypeHip, ::C3, ::UnitRange, ...)    (correlations.jl:201-226)
#   twoPhotons(::SphericalHarmonicsVolume, ::BasisTypeHip, K2, ...)               (correlations.jl:229-243)
#   retrieveSolution_hip(c2, L, LMAX, K2_range, qmax)                             (correlations.jl:262-319)
#   fitStructures_full_hip(volume, reference, num, K_range, L, basis)              (data_processing.jl:26-46)
#   sRAAR_hip(intensity, iterations, beta0, beta_max, tau; cutoff_factor)          (phases.jl:31-95)
#   createSphericalHarmonicsStructure_hip(pdb, LMAX, KMAX, rmax)                    (structure.jl:124-134)
#   rotate_all_at_once_hip(out, params, state, c3ref; checkpoint)                  (determination.jl:281-338)
#   loadHistograms_hip(c3_full, K3)                                                  (datagen.jl:432-475, c3 part)
#
# Written against Julia 0.6 syntax like the rest of the package.

export BasisTypeHip, HIP_init, HIP_calculate_basis, retrieveSolution_hip, fitStructures_full_hip, sRAAR_hip,
       createSphericalHarmonicsStructure_hip, rotate_all_at_once_hip, loadHistograms_hip

const libtp = get(ENV, "THREEPHOTONS_LIBTP", joinpath(dirname(@__FILE__), "..", "libtp.so"))
HIP_enabled = false
hip_ctx = C_NULL

function hip_check(rc::Cint, what::String)
    if rc != 0
        msg = unsafe_string(ccall((:tp_last_error, libtp), Cstring, ()))
        error("$what failed ($rc): $msg")
    end
end

"""Opens device 0 (replaces CUDA_init)."""
function HIP_init(device::Integer=0)
    out = Ref{Ptr{Void}}(C_NULL)
    hip_check(ccall((:tp_init, libtp), Cint, (Cint, Ptr{Ptr{Void}}), device, out), "tp_init")
    global hip_ctx = out[]
    global HIP_enabled = true
    println("Initialization of HIP complete.")
end

"""Precomputed basis resident on the GPU (replaces BasisTypeCuda)."""
type BasisTypeHip <: AbstractBasisType
    handle::Ptr{Void}
    refs::Dict{Any,Ptr{Void}}
    basislen::Int64
    N::Int64
    L::Int64
    LMAX::Int64
    lrange::StepRange
    ctr::Dict
    rtc::Dict
    K::Int64
    lambda::Float64
    dq::Float64
end

function HIP_calculate_basis(L::Int64, LMAX::Int64, N::Int64, K::Int64, lambda::Float64, dq::Float64, forIntensity=true)
    if !HIP_enabled
        return calculate_basis(L, LMAX, N, K, lambda, dq, forIntensity)   # the reference CPU path
    end
    out = Ref{Ptr{Void}}(C_NULL)
    hip_check(ccall((:tp_basis_create, libtp), Cint,
                    (Ptr{Void}, Cint, Cint, Cint, Cint, Cdouble, Cdouble, Cint, Ptr{Ptr{Void}}),
                    hip_ctx, L, LMAX, N, K, lambda, dq, forIntensity ? 1 : 0, out), "tp_basis_create")
    bl = Ref{Int64}(0)
    hip_check(ccall((:tp_basis_info, libtp), Cint, (Ptr{Void}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}),
                    out[], bl, C_NULL, C_NULL, C_NULL), "tp_basis_info")
    lrange = forIntensity ? (0:2:L) : (0:L)
    ctr = Dict(l => Umat(l) for l = lrange)
    rtc = Dict(l => ctranspose(ctr[l]) for l = lrange)
    b = BasisTypeHip(out[], Dict{Any,Ptr{Void}}(), bl[], N, L, LMAX, lrange, ctr, rtc, K, lambda, dq)
    finalizer(b, x -> begin
        for r in values(x.refs) ccall((:tp_ref_destroy, libtp), Void, (Ptr{Void},), r) end
        ccall((:tp_basis_destroy, libtp), Void, (Ptr{Void},), x.handle)
    end)
    return b
end

function HIP_calculate_basis(L::Int64, LMAX::Int64, N::Int64, K::Int64, forIntensity=true)
    return HIP_calculate_basis(L, LMAX, N, K, 0.0, 0.0, forIntensity)
end

global complexBasis_choice = HIP_calculate_basis

"""Coefficients of shells 1..K as one contiguous LMAX^2 x K complex matrix (no conjugation)."""
function hip_pack(intensity::SphericalHarmonicsVolume, K::Int64)
    numcoeff = num_coeff(intensity.LMAX)
    return Complex{Float64}[intensity.coeff[k][i] for i = 1:numcoeff, k = 1:K]
end

function FullCorrelation_parallized(intensity::SphericalHarmonicsVolume, basis::BasisTypeHip, minimal::Bool=true, normalize::Bool=false, return_raw::Bool=false)
    klength = Integer(basis.K*(basis.K+1)*(basis.K+2)/6)
    coeff = hip_pack(intensity, basis.K)
    res = zeros(Float64, 2*basis.N^2, klength)
    hip_check(ccall((:tp_c3, libtp), Cint, (Ptr{Void}, Ptr{Void}, Ptr{Complex{Float64}}, Int64, Ptr{Float64}),
                    hip_ctx, basis.handle, coeff, size(coeff, 1), res), "tp_c3")
    if return_raw return res end
    t = zeros(Float64, basis.N, 2*basis.N, basis.K, basis.K, basis.K)
    i = 1
    for k1=1:basis.K for k2=1:k1 for k3=1:k2
        t[:,:, k3, k2, k1] = reshape(res[:, i], basis.N, 2*basis.N)
        i += 1
    end end end
    return normalize ? t / sum(abs, t) : t
end

"""Device copy of the normalised c3ref slices of the K3_range walk, cached by the
histogram's contents (Base.hash of the array), so an array changed in place gets a
fresh device copy."""
function hip_ref(basis::BasisTypeHip, c3ref::C3, K3_range::UnitRange{Int64})
    key = (hash(c3ref), size(c3ref), K3_range)
    if !haskey(basis.refs, key)
        out = Ref{Ptr{Void}}(C_NULL)
        hip_check(ccall((:tp_ref_create, libtp), Cint, (Ptr{Void}, Ptr{Void}, Ptr{Float64}, Cint, Cint, Cint, Ptr{Ptr{Void}}),
                        hip_ctx, basis.handle, c3ref, size(c3ref, 3), minimum(K3_range), maximum(K3_range), out),
                  "tp_ref_create")
        basis.refs[key] = out[]
    end
    return basis.refs[key]
end

function energy(intensity::SphericalHarmonicsVolume, basis::BasisTypeHip, c3ref::C3, K3_range::UnitRange{Int64}, measure::String="Bayes", negativity_factor::Float64=0.0)
    ref = hip_ref(basis, c3ref, K3_range)
    coeff = hip_pack(intensity, basis.K)
    E = Ref{Float64}(0.0)
    hip_check(ccall((:tp_energy, libtp), Cint, (Ptr{Void}, Ptr{Void}, Ptr{Void}, Ptr{Complex{Float64}}, Int64, Cint, Ptr{Float64}),
                    hip_ctx, basis.handle, ref, coeff, size(coeff, 1), measure == "Bayes" ? 0 : 1, E), "tp_energy")
    return E[]
end

function energy(intensity::SphericalHarmonicsVolume, basis::BasisTypeHip, c3ref::C3, measure::String="Bayes", negativity_factor::Float64=0.0)
    return energy(intensity, basis, c3ref, 1:basis.K, measure, negativity_factor)
end

"""The Bayes energy (FP64) and its gradient: g[:, k] holds dE/dRe c_k + i dE/dIm c_k
for shells k = 1..K in the seanindex order of intensity.coeff[k] (tp_energy_grad; no
counterpart in the reference, which evaluates energy() only)."""
function hip_energy_grad(intensity::SphericalHarmonicsVolume, basis::BasisTypeHip, c3ref::C3, K3_range::UnitRange{Int64}=1:basis.K)
    ref = hip_ref(basis, c3ref, K3_range)
    coeff = hip_pack(intensity, basis.K)
    E = Ref{Float64}(0.0)
    g = zeros(Complex{Float64}, size(coeff))
    hip_check(ccall((:tp_energy_grad, libtp), Cint,
                    (Ptr{Void}, Ptr{Void}, Ptr{Void}, Ptr{Complex{Float64}}, Int64, Ptr{Float64}, Ptr{Complex{Float64}}),
                    hip_ctx, basis.handle, ref, coeff, size(coeff, 1), E, g), "tp_energy_grad")
    return E[], g
end

"""histogramCorrelationsInPicture_alltoall (datagen.jl:228-282) for many pictures at once on the
GPU (tp_hist_*): the pictures' pair and triplet counts are added to c1, c2, c3 with the reference's
weights.  lambda is unused there as in the reference."""
function histogramCorrelationsInPicture_alltoall_hip(pictures::Vector{Vector{Vector{Float64}}}, c1::C1, c2::C2, c3::C3,
                                                      dq::Float64, N::Int64, K2::Int64, K3::Int64, lambda::Float64=0.0)
    h = Ref{Ptr{Void}}(C_NULL)
    hip_check(ccall((:tp_hist_create, libtp), Cint, (Ptr{Void}, Cint, Cint, Cint, Cdouble, Ptr{Ptr{Void}}),
                    hip_ctx, N, K2, K3, dq, h), "tp_hist_create")
    try
        n = sum(length, pictures)
        ph = zeros(Float64, 3, n)
        off = zeros(Int64, length(pictures) + 1)
        i = 0
        for (a, pic) in enumerate(pictures)
            for p in pic
                i += 1
                ph[:, i] = p
            end
            off[a + 1] = i
        end
        hip_check(ccall((:tp_hist_add, libtp), Cint, (Ptr{Void}, Ptr{Float64}, Ptr{Int64}, Cint),
                        h[], ph, off, length(pictures)), "tp_hist_add")
        d1 = zeros(K2); d2 = zeros(N, K2, K2); d3 = zeros(N, 2*N, K3, K3, K3)
        hip_check(ccall((:tp_hist_read, libtp), Cint,
                        (Ptr{Void}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{UInt64}, Ptr{UInt64}, Ptr{UInt64}),
                        h[], d1, d2, d3, C_NULL, C_NULL, C_NULL), "tp_hist_read")
        c1 .+= d1; c2 .+= d2; c3 .+= d3
    finally
        ccall((:tp_hist_destroy, libtp), Void, (Ptr{Void},), h[])
    end
    return nothing
end

"""twoPhotons (correlations.jl:229-243) on the GPU: c2[:, k2, k1] of shells 1:K2 over the basis's
l range and alpha grid (lambda = 0)."""
function twoPhotons(volume::SphericalHarmonicsVolume, basis::BasisTypeHip, K2::Int64, minimal::Bool=false, normalize::Bool=false)
    coeff = hip_pack(volume, K2)
    c = zeros(Float64, basis.N, K2, K2)
    # alpha_star with basis.lambda and dq(volume), as the reference's twoPhotons (correlations.jl:231,238)
    hip_check(ccall((:tp_two_photons_ewald, libtp), Cint,
                    (Ptr{Void}, Ptr{Void}, Ptr{Complex{Float64}}, Int64, Cint, Cint, Cint, Cdouble, Ptr{Float64}),
                    hip_ctx, basis.handle, coeff, size(coeff, 1), K2, minimal, normalize, dq(volume), c),
              "tp_two_photons_ewald")
    return c
end

"""retrieveSolution (correlations.jl:262-319) through libtp (host code, any lambda):
the start structure whose two-photon correlation matches c2 on K2_range."""
function retrieveSolution_hip(c2::C2, L::Int64, LMAX::Int64, K2_range::UnitRange{Int64}, qmax::Float64, lambda::Float64=0.0)
    N, K2dim, _ = size(c2)
    hi = maximum(K2_range)
    coeff = zeros(Complex{Float64}, LMAX^2, hi)
    flipped = Ref{Cint}(0); neg = Ref{Cdouble}(0.0)
    hip_check(ccall((:tp_retrieve_solution_ewald, libtp), Cint,
                    (Ptr{Float64}, Cint, Cint, Cint, Cint, Cint, Cint, Cdouble, Cdouble, Ptr{Complex{Float64}}, Int64,
                     Ptr{Cint}, Ptr{Cdouble}),
                    Array{Float64}(c2), N, K2dim, minimum(K2_range), hi, L, LMAX, qmax, lambda, coeff, LMAX^2,
                    flipped, neg), "tp_retrieve_solution_ewald")
    intensity = SphericalHarmonicsVolume(LMAX, hi, qmax)
    for k = 1:hi intensity.coeff[k][:] = coeff[:, k] end
    return intensity
end

"""fitStructures_full (data_processing.jl:26-46) with every candidate rotation's similarity
computed by libtp in one call (tp_fit_*).  basis: a BasisTypeHip with the same L and LMAX."""
function fitStructures_full_hip(volume::SphericalHarmonicsVolume, reference::SphericalHarmonicsVolume, num::Int64,
                                K_range::UnitRange{Int64}, L::Int64, basis::BasisTypeHip, ta::Float64=0.0,
                                tb::Float64=float(pi)/2, pa::Float64=0.0, pb::Float64=float(pi), ga::Float64=0.0,
                                gb::Float64=float(pi))
    K = maximum(K_range)
    angles = [(theta, phi, gamma) for theta = ta:(tb-ta)/num:tb for phi = pa:(pb-pa)/num:pb for gamma = ga:(gb-ga)/num:gb]
    euler = Float64[a[i] for i = 1:3, a in angles]
    f = Ref{Ptr{Void}}(C_NULL)
    hip_check(ccall((:tp_fit_create, libtp), Cint, (Ptr{Void}, Ptr{Void}, Ptr{Complex{Float64}}, Int64, Cint, Cint, Ptr{Ptr{Void}}),
                    hip_ctx, basis.handle, hip_pack(reference, K), num_coeff(reference.LMAX), minimum(K_range), K, f),
              "tp_fit_create")
    sc = zeros(Float64, length(angles))
    try
        hip_check(ccall((:tp_fit_set_volume, libtp), Cint, (Ptr{Void}, Ptr{Complex{Float64}}, Int64),
                        f[], hip_pack(volume, K), num_coeff(volume.LMAX)), "tp_fit_set_volume")
        hip_check(ccall((:tp_fit_similarity, libtp), Cint, (Ptr{Void}, Cint, Ptr{Float64}, Ptr{Float64}),
                        f[], length(angles), euler, sc), "tp_fit_similarity")
    finally
        ccall((:tp_fit_destroy, libtp), Void, (Ptr{Void},), f[])
    end
    i = indmax(sc)
    bestt, bestp, bestg = angles[i]
    return (rotateStructure(volume, bestt, bestp, bestg, K, 2:2:L), sc[i], bestt, bestp, bestg)
end

"""sRAAR (phases.jl:31-95) through libtp: every transform, projection and reflection on the GPU.
phases: initial phases ((2 LMAX)^2 x KMAX, surface sample order) or nothing (drawn from `seed`)."""
function sRAAR_hip(intensity::SphericalHarmonicsVolume, iterations::Integer, beta0::Float64=0.75,
                   beta_max::Float64=0.75, tau::Float64=100.0; cutoff_factor::Float64=0.5,
                   phases::Union{Void, Matrix{Float64}}=nothing, seed::Integer=0)
    K, LMAX = intensity.KMAX, intensity.LMAX
    coeff = hip_pack(intensity, K)
    out = zeros(Complex{Float64}, size(coeff))
    ph = phases == nothing ? C_NULL : phases
    hip_check(ccall((:tp_sraar, libtp), Cint,
                    (Ptr{Void}, Cint, Cint, Ptr{Complex{Float64}}, Int64, Cint, Cdouble, Cdouble, Cdouble, Cdouble,
                     Ptr{Float64}, UInt64, Ptr{Complex{Float64}}, Ptr{Float64}),
                    hip_ctx, LMAX, K, coeff, size(coeff, 1), iterations, beta0, beta_max, tau, cutoff_factor,
                    ph, UInt64(seed), out, C_NULL), "tp_sraar")
    density = SphericalHarmonicsVolume(LMAX, K, qmax(intensity))
    for k = 1:K density.coeff[k][:] = out[:, k] end
    return density
end

"""Switch a basis's arithmetic (TP_PREC_* of include/tp.h): 0 = FP64, 1 = FP16 hi/lo
triple products and contraction, 2 = BF16, 3 = FP16 hi/lo triple products with the FP64
contraction, 4 = FP64 triple products with the FP16 hi/lo contraction."""
function hip_set_precision(basis::BasisTypeHip, mode::Integer)
    hip_check(ccall((:tp_basis_set_precision, libtp), Cint, (Ptr{Void}, Cint), basis.handle, mode),
              "tp_basis_set_precision")
end

"""Many rotate_all_at_once runs in one process: run g starts from intensities[g],
scores against c3refs[g] over shells 1:size(c3refs[g], 3) on `basis` (mixed precision,
K >= every run's), and every sweep evaluates `nprop` proposals per run in one graph launch.
Returns the final energies; the runs' structures are read back into `intensities`."""
function rotate_all_at_once_batched(intensities::Vector{SphericalHarmonicsVolume}, basis::BasisTypeHip,
                                    c3refs::Vector{C3}, seeds::Vector{UInt64}, nsweeps::Int;
                                    nprop::Int=64, temperature_factor::Float64=0.1,
                                    decay::Float64=0.99998, factor::Float64=1.05)
    n = length(intensities)
    refs = Ptr{Void}[hip_ref(basis, c3refs[g], 1:size(c3refs[g], 3)) for g in 1:n]
    coeffs = [hip_pack(intensities[g], size(c3refs[g], 3)) for g in 1:n]
    ptrs = Ptr{Complex{Float64}}[pointer(c) for c in coeffs]
    lds = Int64[size(c, 1) for c in coeffs]
    batch = Ref{Ptr{Void}}(C_NULL)
    hip_check(ccall((:tp_batch_create, libtp), Cint,
                    (Ptr{Void}, Ptr{Void}, Cint, Cint, Ptr{Ptr{Void}}, Ptr{Ptr{Complex{Float64}}}, Ptr{Int64},
                     Ptr{UInt64}, Cint, Ptr{Ptr{Void}}),
                    hip_ctx, basis.handle, n, nprop, refs, ptrs, lds, seeds, nsweeps, batch), "tp_batch_create")
    try
        hip_check(ccall((:tp_batch_probe, libtp), Cint, (Ptr{Void}, Cdouble), batch[], temperature_factor),
                  "tp_batch_probe")
        hip_check(ccall((:tp_batch_run, libtp), Cint, (Ptr{Void}, Cint, Cdouble, Cdouble), batch[], nsweeps, decay,
                        factor), "tp_batch_run")
        E = zeros(n)
        for g in 1:n
            Eg = Ref{Float64}(0.0)
            hip_check(ccall((:tp_batch_group, libtp), Cint, (Ptr{Void}, Cint, Ptr{Float64}, Ptr{Float64},
                            Ptr{Float64}, Ptr{Int32}, Ptr{Int32}), batch[], g - 1, Eg, C_NULL, C_NULL, C_NULL,
                            C_NULL), "tp_batch_group")
            E[g] = Eg[]
            out = copy(coeffs[g])
            hip_check(ccall((:tp_batch_coeff, libtp), Cint, (Ptr{Void}, Cint, Ptr{Complex{Float64}}, Int64),
                            batch[], g - 1, out, size(out, 1)), "tp_batch_coeff")
            for k in 1:size(out, 2)
                intensities[g].coeff[k][:] = out[:, k]
            end
        end
        return E
    finally
        ccall((:tp_batch_destroy, libtp), Void, (Ptr{Void},), batch[])
    end
end

"""createSphericalHarmonicsStructure (structure.jl:124-134) with getDensity,
forward and absoluteSquare on the device: (density, fourier, intensity)."""
function createSphericalHarmonicsStructure_hip(moleculepdb, LMAX::Int64, KMAX::Int64, rmax::Float64)
    rows = filter(atom -> atom.t != "1", loadPDB(moleculepdb))
    atoms = zeros(Float64, 5, length(rows))
    for (i, atom) in enumerate(rows)
        ff = atomfactorlist[atom.t]
        atoms[:, i] = [atom.pos[1], atom.pos[2], atom.pos[3], ff.R, ff.B]
    end
    n = num_coeff(LMAX)
    d = zeros(Complex{Float64}, n, KMAX)
    f = zeros(Complex{Float64}, n, KMAX)
    I = zeros(Complex{Float64}, n, KMAX)
    hip_check(ccall((:tp_structure_from_atoms, libtp), Cint,
                    (Ptr{Void}, Ptr{Float64}, Cint, Cint, Cint, Cdouble, Ptr{Complex{Float64}},
                     Ptr{Complex{Float64}}, Ptr{Complex{Float64}}, Int64),
                    hip_ctx, atoms, size(atoms, 2), LMAX, KMAX, rmax, d, f, I, n), "tp_structure_from_atoms")
    density = SphericalHarmonicsVolume(LMAX, KMAX, rmax)
    fourier = SphericalHarmonicsVolume(LMAX, KMAX, qmax(KMAX, rmax))
    intensity = SphericalHarmonicsVolume(LMAX, KMAX, qmax(KMAX, rmax))
    for k = 1:KMAX
        density.coeff[k][:] = d[:, k]
        fourier.coeff[k][:] = f[:, k]
        intensity.coeff[k][:] = I[:, k]
    end
    return density, fourier, intensity
end

# isbits mirrors of tp_anneal_params, tp_anneal_status and tp_checkpoint_info (include/tp.h)
immutable TpAnnealParams
    nsteps::Cint
    ntemp::Cint
    initial_stepsize::Cdouble
    stepsizefactor::Cdouble
    initial_temperature_factor::Cdouble
    temperature_decay::Cdouble
    seed::UInt64
end

immutable TpAnnealStatus
    step::Int64
    E::Cdouble
    T::Cdouble
    E0::Cdouble
    T0::Cdouble
    stepsizes::NTuple{32,Cdouble}
    nl::Int32
    naccepted::Int32
    stopped::Int32
    reserved::Int32
end

immutable TpCheckpointInfo
    params::TpAnnealParams
    status::TpAnnealStatus
    min_stepsize::Cdouble
    lambda::Cdouble
    dq::Cdouble
    L::Int32
    LMAX::Int32
    N::Int32
    K::Int32
    k3lo::Int32
    k3hi::Int32
    nshell::Int32
    ld::Int32
end

"""rotate_all_at_once (determination.jl:281-338) with every step on the GPU: the
intensity stays on the device between steps (tp_state), the steps run in pieces of
`chunk` (tp_anneal_steps) until stepsizes[L] <= minstepsize (determination.jl:1, 322),
and state["i"], state["E"], state["T"], state["stepsizes"] and state["intensity"] are
written back after every piece as the reference keeps them.  regular_action's save
every 500 steps (determination.jl:270-278) becomes a checkpoint file at `checkpoint`
(tp_checkpoint_save; nothing: none).  A state that already has stepsizes for this L
continues from its i, E, T and stepsizes, as the reference's continuation does."""
function rotate_all_at_once_hip(out, params::Dict, state::Dict, c3ref::C3; checkpoint=nothing, chunk::Int=500)
    state["L"] = params["L"]
    L = state["L"]
    basis = HIP_calculate_basis(L, params["LMAX"], params["N"], maximum(params["K3_range"]), params["lambda"],
                                dq(state["intensity"]))
    ref = hip_ref(basis, c3ref, params["K3_range"])
    intensity = state["intensity"]
    nshell = intensity.KMAX
    coeff = hip_pack(intensity, nshell)
    s = Ref{Ptr{Void}}(C_NULL)
    hip_check(ccall((:tp_state_create, libtp), Cint,
                    (Ptr{Void}, Ptr{Void}, Ptr{Complex{Float64}}, Int64, Cint, Ptr{Ptr{Void}}),
                    hip_ctx, basis.handle, coeff, size(coeff, 1), nshell, s), "tp_state_create")
    p = Ref(TpAnnealParams(0, 50, params["initial_stepsize"], params["stepsizefactor"],
                           params["initial_temperature_factor"], params["temperature_decay"],
                           UInt64(get(params, "seed", 0))))
    nl = div(L, 2)
    st = Ref{TpAnnealStatus}()
    try
        if !haskey(state["stepsizes"], L)
            state["K"] = maximum(params["K3_range"])
            hip_check(ccall((:tp_anneal_begin, libtp), Cint,
                            (Ptr{Void}, Ptr{Void}, Ptr{Void}, Ptr{Void}, Ref{TpAnnealParams}, Ref{TpAnnealStatus}),
                            hip_ctx, basis.handle, ref, s[], p, st), "tp_anneal_begin")
        else
            steps = ntuple(i -> i <= nl ? Float64(state["stepsizes"][2 * i]) : 0.0, 32)
            st[] = TpAnnealStatus(state["i"], state["E"], state["T"], state["E"], state["T"], steps, nl, 0, 0, 0)
        end
        while true
            done = Ref{Int32}(0)
            hip_check(ccall((:tp_anneal_steps, libtp), Cint,
                            (Ptr{Void}, Ptr{Void}, Ptr{Void}, Ptr{Void}, Ref{TpAnnealParams}, Ref{TpAnnealStatus},
                             Cint, Cdouble, Ptr{Float64}, Ptr{Float64}, Ptr{Int32}, Ref{Int32}),
                            hip_ctx, basis.handle, ref, s[], p, st, chunk, minstepsize, C_NULL, C_NULL, C_NULL, done),
                      "tp_anneal_steps")
            hip_check(ccall((:tp_state_read, libtp), Cint, (Ptr{Void}, Ptr{Void}, Ptr{Complex{Float64}}, Int64),
                            hip_ctx, s[], coeff, size(coeff, 1)), "tp_state_read")
            for k = 1:nshell intensity.coeff[k][:] = coeff[:, k] end
            state["i"], state["E"], state["T"] = st[].step, st[].E, st[].T
            state["stepsizes"] = Dict(2 * i => st[].stepsizes[i] for i = 1:nl)
            if checkpoint != nothing
                info = Ref(TpCheckpointInfo(p[], st[], minstepsize, params["lambda"], dq(intensity), L, params["LMAX"],
                                            params["N"], maximum(params["K3_range"]), minimum(params["K3_range"]),
                                            maximum(params["K3_range"]), nshell, size(coeff, 1)))
                hip_check(ccall((:tp_checkpoint_save, libtp), Cint, (Cstring, Ref{TpCheckpointInfo}, Ptr{Complex{Float64}}),
                                checkpoint, info, coeff), "tp_checkpoint_save")
            end
            println(out, "step $(state["i"]) E $(state["E"]) T $(state["T"]) stepsize $(st[].stepsizes[nl])")
            flush(out)
            if st[].stopped != 0 || done[] == 0
                break
            end
        end
    finally
        ccall((:tp_state_destroy, libtp), Void, (Ptr{Void},), s[])
    end
    return state
end

"""The c3 part of loadHistograms (datagen.jl:432-475) through libtp: an old N x N
histogram is doubled along beta, entries are clamped to >= 1e-30, and the first K3
shells are kept and divided by their total."""
function loadHistograms_hip(c3_full::C3, K3::Int64)
    NA, NB, Kf = size(c3_full, 1), size(c3_full, 2), size(c3_full, 3)
    N = NA
    out = zeros(Float64, N, 2 * N, K3, K3, K3)
    hip_check(ccall((:tp_load_histograms, libtp), Cint, (Ptr{Float64}, Cint, Cint, Cint, Cint, Ptr{Float64}),
                    Array{Float64}(c3_full), NA, NB, Kf, K3, out), "tp_load_histograms")
    return out
end
