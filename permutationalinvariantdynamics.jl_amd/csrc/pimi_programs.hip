// Compiled banded programs: for a stencil on the banded form, the chain and
// entry structure of every output sector (chain count, strip width, and per
// entry its kind, strip column and row shift) is turned into straight-line
// device code at plan set-up and compiled with hiprtc for gfx950.  Sectors
// of one structure share a program; the qubit plans have three (the first
// and last sector and all others).  A compiled program makes the same loads
// and FMAs in the same order as the interpreting kernel, so its results are
// bitwise those of the interpreting kernel; what it removes is the per-entry
// dispatch (descriptor waits, the kind switch, the shift loop), and it lets
// the compiler issue a chain's loads together.  The code depends on the
// structure only (coefficients, offsets and sizes stay in the tables), so
// one module serves every plan of that structure on a device (cached for
// the process).
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "pimi_internal.hpp"
#include "pimi_rtc_headers.inc"  // kRtcBaseHeader, kRtcDiaHeader: pimi_base.hpp and pimi_dia.hpp as text

namespace pimi {

namespace {

constexpr int kMaxProgramEntries = 96;  // longer sector structures stay interpreted

const char *prec_name(int p) { return p == 0 ? "f64" : "f32"; }

std::string kernel_name(const char *kind, int prec, int tb) {
    return std::string("pimi_dia_") + kind + "_" + prec_name(prec) + "_" + std::to_string(tb);
}

// the structure of sector s, as a key: per chain its strip width, then per
// entry kind, strip column and row shift
std::string sector_key(const DevStencil &S, int s) {
    std::string k;
    for (int32_t g = S.h_cbase[(size_t)s]; g < S.h_cbase[(size_t)s + 1]; ++g) {
        const DiaChain &G = S.h_chains[(size_t)g];
        k += G.span == 0 ? 'n' : 'w';
        for (int32_t e = G.e0; e < G.e1; ++e) {
            const DiaEntry &D = S.h_dents[(size_t)e];
            k += (char)('0' + D.kind);
            k += (char)('0' + D.dB);
            k += (char)('0' + D.sh);
        }
        k += ';';
    }
    return k;
}

// straight-line code of one structure: chain k runs on wave k mod SPLIT
std::string program_source(const DevStencil &S, int s, int id) {
    std::ostringstream o;
    o << "template <typename R, int TB, int SPLIT>\n"
         "__device__ __forceinline__ void prog_" << id
      << "(cx<R> (&acc)[TB], const DiaLane<R, TB> &L, const DiaTables &T, int s, int wave) {\n"
         "    const PIMI_CONST DiaChain *C = cptr(T.chains) + cptr(T.cbase)[s];\n"
         "    const PIMI_CONST DiaEntry *E = cptr(T.ents) + C[0].e0;\n";
    const int32_t g0 = S.h_cbase[(size_t)s], g1 = S.h_cbase[(size_t)s + 1];
    const int32_t e_first = g1 > g0 ? S.h_chains[(size_t)g0].e0 : 0;
    for (int32_t g = g0; g < g1; ++g) {
        const DiaChain &G = S.h_chains[(size_t)g];
        o << "    if (" << (g - g0) << " % SPLIT == wave) {\n"
          << "        const DiaChain G = C[" << (g - g0) << "];\n"
          << "        cx<R> st[TB + 2];\n"
          << "        L.template strip<" << (G.span == 0 ? "TB" : "TB + 2") << ">(st, G);\n";
        int shifted = 0;
        for (int32_t e = G.e0; e < G.e1; ++e) {
            const DiaEntry &D = S.h_dents[(size_t)e];
            for (; shifted < D.sh; ++shifted) o << "        dia_shift<R, TB>(st);\n";
            o << "        L.template entry<" << D.kind << ", " << D.dB << ">(acc, st, E[" << (e - e_first)
              << "]);\n";
        }
        o << "    }\n";
    }
    o << "}\n\n";
    return o.str();
}

// the programs and their dispatch on the sector's program id (every sector
// of a stencil with programs has one)
std::string programs_source(const std::vector<std::string> &progs) {
    std::ostringstream o;
    o << "#include \"pimi_dia.hpp\"\nnamespace pimi {\n";
    for (const std::string &p : progs) o << p;
    o << "struct DiaCompiled {\n"
         "    template <typename R, int TB, int SPLIT>\n"
         "    static __device__ __forceinline__ void run(cx<R> (&acc)[TB], const DiaLane<R, TB> &L, "
         "const DiaTables &T, int s, int wave) {\n"
         "        switch (cptr(T.pid)[s]) {\n";
    for (size_t k = 0; k < progs.size(); ++k)
        o << "            case " << k << ": prog_" << k << "<R, TB, SPLIT>(acc, L, T, s, wave); break;\n";
    o << "            default: break;\n"
         "        }\n"
         "    }\n"
         "};\n"
         "}  // namespace pimi\n";
    return o.str();
}

// one kernel over the programs: form 0 the single-vector (split) apply,
// 1 the batched one, in precision prec (0 FP64, 1 FP32), tile width tb
std::string kernel_source(int form, int prec, int tb) {
    std::ostringstream o;
    const char *R = prec == 0 ? "double" : "float";
    if (form == 0)
        o << "extern \"C\" __global__ __launch_bounds__(pimi::kApplyWave * pimi::kDiaSplit) void "
          << kernel_name("split", prec, tb) << "(const pimi::cx<" << R << "> *x, pimi::cx<" << R
          << "> *y, int64_t n, const pimi::ApplyTile *tiles, int64_t ntiles, int64_t bpx, const int64_t *offs, "
             "const int32_t *dims, pimi::DiaTables T, const pimi::cx<"
          << R << "> *coef, " << R << " out_scale, int64_t ncols) {\n"
          << "    pimi::dia_split_body<pimi::DiaCompiled, " << R << ", " << tb
          << ">(x, y, n, tiles, ntiles, bpx, offs, dims, T, coef, out_scale, ncols);\n}\n";
    else
        o << "extern \"C\" __global__ __launch_bounds__(pimi::kApplyWave * pimi::kDiaWaves) void "
          << kernel_name("batch", prec, tb) << "(const pimi::BDia *J, int nj, int64_t total, int64_t per_xcd) {\n"
          << "    pimi::dia_batch_body<pimi::DiaCompiled, " << R << ", " << tb << ">(J, nj, total, per_xcd);\n}\n";
    return o.str();
}

// the gfx950 code object of a module source (hiprtc; needs no device)
std::vector<char> compile_code(const std::string &src) {
    hiprtcProgram prog;
    const char *hdr_src[2] = {kRtcBaseHeader, kRtcDiaHeader};
    const char *hdr_name[2] = {"pimi_base.hpp", "pimi_dia.hpp"};
    if (hiprtcCreateProgram(&prog, src.c_str(), "pimi_dia_programs.hip", 2, hdr_src, hdr_name) != HIPRTC_SUCCESS)
        throw Error(PIMI_EDEVICE, "hiprtc: cannot create the banded program module");
    // no packed FP32 instructions, as for every other kernel of the library
    const char *opts[] = {"--gpu-architecture=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
                          "-packed-fp32-ops", "-mllvm", "-structurizecfg-skip-uniform-regions=true"};
    const hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
    if (rc != HIPRTC_SUCCESS) {
        size_t ls = 0;
        hiprtcGetProgramLogSize(prog, &ls);
        std::string log(ls, '\0');
        if (ls) hiprtcGetProgramLog(prog, &log[0]);
        hiprtcDestroyProgram(&prog);
        throw Error(PIMI_EDEVICE, "hiprtc: banded program module failed to compile: " + log.substr(0, 2000));
    }
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    std::vector<char> code(cs);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    return code;
}

// the programs of stencil S: per-sector ids, the module source, and the
// structure signature (the module cache key)
struct Programs {
    std::vector<int32_t> pid;
    std::string source, signature;
    int count = 0;
};

Programs make_programs(const DevStencil &S) {
    Programs out;
    const int ns = (int)S.h_cbase.size() - 1;
    std::map<std::string, int> ids;
    std::vector<std::string> progs;
    out.pid.assign((size_t)ns, -1);
    for (int s = 0; s < ns; ++s) {
        const int32_t g0 = S.h_cbase[(size_t)s], g1 = S.h_cbase[(size_t)s + 1];
        const int32_t count = g1 > g0 ? S.h_chains[(size_t)g1 - 1].e1 - S.h_chains[(size_t)g0].e0 : 0;
        if (count > kMaxProgramEntries) return Programs{};  // too long: the stencil stays interpreted
        const std::string key = sector_key(S, s);
        auto it = ids.find(key);
        if (it == ids.end()) {
            it = ids.emplace(key, (int)progs.size()).first;
            progs.push_back(program_source(S, s, it->second));
        }
        out.pid[(size_t)s] = it->second;
    }
    for (auto &kv : ids) out.signature += kv.first + "|" + std::to_string(kv.second) + "\n";
    out.source = programs_source(progs);
    out.count = (int)progs.size();
    return out;
}

}  // namespace

// PIMI_DIA_PROGRAMS=0 keeps every stencil on the interpreting kernels
static bool dia_programs_enabled() {
    const char *v = std::getenv("PIMI_DIA_PROGRAMS");
    return !v || std::atoi(v) != 0;
}

void dia_setup_programs(DevStencil &S, int device) {
    if (!S.dia_ok || !dia_programs_enabled()) return;
    Programs G = make_programs(S);
    if (G.signature.empty()) return;
    S.h_pid = G.pid;
    S.pid.upload(S.h_pid.data(), S.h_pid.size());
    S.prog_src = std::move(G.source);
    S.prog_device = device;
    S.prog_sig = std::move(G.signature);  // set last: a non-empty signature means ready
}

hipFunction_t dia_program(const DevStencil &S, int form, int prec, int w, hipStream_t stream) {
    if (S.prog_sig.empty()) return nullptr;
    hipFunction_t f = S.prog_fn[form][prec][w].load(std::memory_order_acquire);
    if (f) return f;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return nullptr;
    // kernels of one structure, form, precision and width are compiled once
    // per device for the process; modules stay loaded until it ends
    static std::mutex mu;
    static std::map<std::string, hipFunction_t> cache;
    const std::string key = std::to_string(S.prog_device) + "/" + std::to_string(form) + "/" + std::to_string(prec) +
                            "/" + std::to_string(w) + "\n" + S.prog_sig;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it == cache.end()) {
        const std::vector<char> code = compile_code(S.prog_src + kernel_source(form, prec, kDiaWidths[w]));
        int cur = 0;
        HIPCHK(hipGetDevice(&cur));
        if (cur != S.prog_device) HIPCHK(hipSetDevice(S.prog_device));
        hipModule_t mod = nullptr;
        hipFunction_t fn = nullptr;
        HIPCHK(hipModuleLoadData(&mod, code.data()));
        HIPCHK(hipModuleGetFunction(&fn, mod, kernel_name(form == 0 ? "split" : "batch", prec, kDiaWidths[w]).c_str()));
        if (cur != S.prog_device) HIPCHK(hipSetDevice(cur));
        it = cache.emplace(key, fn).first;
    }
    S.prog_fn[form][prec][w].store(it->second, std::memory_order_release);
    return it->second;
}

}  // namespace pimi

extern "C" pimi_status pimi_plan_banded_programs(const pimi_plan *P, int32_t compile, char *source, int64_t cap,
                                                 int64_t *len, int32_t *nprograms, int32_t *pid) {
    return pimi::guarded([&] {
        if (!P || !len || !nprograms) throw pimi::Error(PIMI_EINVAL, "null argument");
        const pimi::DevStencil &S = P->main;
        if (!S.dia_ok) {
            *len = 0;
            *nprograms = 0;
            return;
        }
        pimi::Programs G = pimi::make_programs(S);
        *len = (int64_t)G.source.size();
        *nprograms = G.count;
        if (source && cap > 0) {
            const size_t k = (size_t)std::min<int64_t>(cap - 1, (int64_t)G.source.size());
            std::copy(G.source.begin(), G.source.begin() + (ptrdiff_t)k, source);
            source[k] = '\0';
        }
        if (pid) std::copy(G.pid.begin(), G.pid.end(), pid);
        if (compile && G.count)
            for (int form = 0; form < 2; ++form) (void)pimi::compile_code(G.source + pimi::kernel_source(form, 1, 8));
    });
}
