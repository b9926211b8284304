// The C++ drop-in header on the device: gfa::action_t::multiply through gfa_action_multiply with the plan's
// colindx in device memory (as tfQMRgpu passes it), gfa::green_solver_t::solve for the action keys 64021 and
// 32021, the refused keys 64022/32022 and a wrong colIndex.  Inputs come from files a test wrote (directory
// argv[1]); the outputs are written there for the test to compare with the oracle:
//   in:  desc.txt (ng[3] bc[3] h r_trunc E_re E_im k[3]), xyzZinso.bin, V.bin, hs.bin (per atom: int32 nc,
//        double[2][nc][nc]), x.bin (double[nnzb][2][64][64])
//   out: y64.bin (double), y32.bin (float), rho64.bin, rho32.bin (complex double [nCols][64]), info.txt
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "gfa_action.hxx"

namespace {

template <typename T>
std::vector<T> read_all(std::string const& path) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) { std::fprintf(stderr, "cannot read %s\n", path.c_str()); std::exit(3); }
    size_t const n = size_t(f.tellg());
    f.seekg(0);
    std::vector<T> v(n / sizeof(T));
    f.read(reinterpret_cast<char*>(v.data()), std::streamsize(v.size() * sizeof(T)));
    return v;
}

template <typename T>
void write_all(std::string const& path, T const* p, size_t n) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<char const*>(p), std::streamsize(n * sizeof(T)));
}

#define HIPCHECK(e) do { hipError_t const r_ = (e); if (r_ != hipSuccess) { \
    std::fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(r_), __LINE__); std::exit(4); } } while (0)

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) { std::fprintf(stderr, "usage: drop_in_gpu <dir>\n"); return 2; }
    std::string const dir = argv[1];
    uint32_t ng[3];
    int bci[3];
    double h, r_trunc, Er, Ei, k[3];
    {
        std::ifstream f(dir + "/desc.txt");
        f >> ng[0] >> ng[1] >> ng[2] >> bci[0] >> bci[1] >> bci[2] >> h >> r_trunc >> Er >> Ei >> k[0] >> k[1] >> k[2];
        if (!f) { std::fprintf(stderr, "bad desc.txt\n"); return 3; }
    }
    int8_t const bc[3] = {int8_t(bci[0]), int8_t(bci[1]), int8_t(bci[2])};
    double const hg[3] = {h, h, h};
    int const fd[3] = {4, 4, 4};
    auto const xyz = read_all<double>(dir + "/xyzZinso.bin");
    auto const V = read_all<double>(dir + "/V.bin");
    std::vector<std::vector<double>> hs;
    {
        auto const raw = read_all<char>(dir + "/hs.bin");
        size_t off = 0;
        while (off < raw.size()) {
            int32_t nc;
            std::memcpy(&nc, raw.data() + off, 4);
            off += 4;
            std::vector<double> m(size_t(2) * nc * nc);
            std::memcpy(m.data(), raw.data() + off, m.size() * 8);
            off += m.size() * 8;
            hs.push_back(std::move(m));
        }
    }
    std::FILE* info = std::fopen((dir + "/info.txt").c_str(), "w");

    gfa::plan_t plan(ng, bc, hg, xyz, r_trunc, fd);
    plan.update_potential(V.data(), hs);
    plan.update_energy_parameter({Er, Ei});
    plan.update_phases(k);
    uint64_t const nnzb = plan.colindx.size();
    std::fprintf(info, "nnzb %llu nRows %u nCols %u rowstart_back %u subset0 %u\n", (unsigned long long)nnzb,
                 plan.nRows, plan.nCols, plan.rowstart.back(), plan.subset.at(0));

    // x in both precisions, the plan's colindx in device memory
    auto const x64 = read_all<double>(dir + "/x.bin");
    size_t const n = size_t(nnzb) * 2 * 64 * 64;
    if (x64.size() != n) { std::fprintf(stderr, "x.bin has %zu values, expected %zu\n", x64.size(), n); return 3; }
    std::vector<float> x32(x64.begin(), x64.end());
    double *dx64, *dy64;
    float *dx32, *dy32;
    uint16_t *dcol, *dbad;
    HIPCHECK(hipMalloc(&dx64, n * 8));
    HIPCHECK(hipMalloc(&dy64, n * 8));
    HIPCHECK(hipMalloc(&dx32, n * 4));
    HIPCHECK(hipMalloc(&dy32, n * 4));
    HIPCHECK(hipMalloc(&dcol, nnzb * 2));
    HIPCHECK(hipMalloc(&dbad, nnzb * 2));
    HIPCHECK(hipMemcpy(dx64, x64.data(), n * 8, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(dx32, x32.data(), n * 4, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(dcol, plan.colindx.data(), nnzb * 2, hipMemcpyHostToDevice));
    std::vector<uint16_t> bad(plan.colindx);
    bad.back() = uint16_t(bad.back() + 1);
    HIPCHECK(hipMemcpy(dbad, bad.data(), nnzb * 2, hipMemcpyHostToDevice));

    typedef gfa::action_t<double, 2, 1, 64> Act821;
    typedef gfa::action_t<float, 2, 1, 64> Act421;
    typedef gfa::action_t<double, 2, 2, 64> Act822;
    static_assert(Act821::LM == 64 && Act821::LN == 64 && Act822::LM == 128, "block sizes");
    Act821 a64(&plan);
    Act421 a32(&plan);
    double const f64 = a64.multiply((double(*)[2][64][64])dy64, (double const(*)[2][64][64])dx64, dcol, uint32_t(nnzb),
                                    plan.nCols, 0, nullptr, false);
    double const f32 = a32.multiply((float(*)[2][64][64])dy32, (float const(*)[2][64][64])dx32, dcol, uint32_t(nnzb),
                                    plan.nCols, 0, nullptr, false);
    HIPCHECK(hipDeviceSynchronize());
    std::vector<double> y64(n);
    std::vector<float> y32(n);
    HIPCHECK(hipMemcpy(y64.data(), dy64, n * 8, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(y32.data(), dy32, n * 4, hipMemcpyDeviceToHost));
    write_all(dir + "/y64.bin", y64.data(), n);
    write_all(dir + "/y32.bin", y32.data(), n);
    std::fprintf(info, "flops64 %.6e flops32 %.6e has_preconditioner %d\n", f64, f32, int(a64.has_preconditioner()));

    // refused: a colIndex that differs from the plan's, and the non-collinear action
    try {
        a64.multiply((double(*)[2][64][64])dy64, (double const(*)[2][64][64])dx64, dbad, uint32_t(nnzb), plan.nCols);
        std::fprintf(info, "bad_colindex accepted\n");
    } catch (gfa::error const& e) {
        std::fprintf(info, "bad_colindex status %d\n", e.status);
    }
    try {
        Act822 a822(&plan);
        a822.multiply(nullptr, nullptr, dcol, uint32_t(nnzb), plan.nCols);
        std::fprintf(info, "noco2_multiply accepted\n");
    } catch (gfa::error const& e) {
        std::fprintf(info, "noco2_multiply status %d\n", e.status);
    }

    // the solver facade: keys 64021 and 32021 solve, 64022 and 32022 are refused with a status
    std::vector<std::complex<double>> rho(size_t(plan.nCols) * 64);
    for (int fp : {64, 32}) {
        gfa::green_solver_t solver(&plan, 0, 0, fp, 1);
        gfa_solve_stats st{};
        int const s = solver.solve(rho.data(), plan.nCols, 2000, 1, 0, &st);
        std::fprintf(info, "key %d status %d converged %d of %d max_residual %.3e iterations_needed %d "
                           "residuum_reached %.3e gpu_mem %zu\n", solver.action_key(), s, st.converged_columns,
                     st.total_columns, st.max_residual, plan.iterations_needed, double(plan.residuum_reached),
                     plan.gpu_mem);
        write_all(dir + (fp == 64 ? "/rho64.bin" : "/rho32.bin"), rho.data(), rho.size());
        gfa::green_solver_t noco(&plan, 0, 0, fp, 2);
        std::fprintf(info, "key %d status %d\n", noco.action_key(), noco.solve(rho.data(), plan.nCols, 2000));
    }
    {   // key 32021 with the mixed COCG policy, set through the solver's public members: one FP32 solve with
        // residual replacement, groups still above 1e-3 after 8 FP32 applications go on in FP64
        gfa::green_solver_t solver(&plan, 0, 0, 32, 1);
        solver.method = 1;
        solver.inner_tolerance = 1e-30;
        solver.max_refinements = 6;
        solver.fp32_max_iterations = 1200;
        solver.fp32_progress_check = 8;
        gfa_solve_stats st{};
        int const s = solver.solve(rho.data(), plan.nCols, 2000, 1, 0, &st);
        std::fprintf(info, "cocg32 status %d converged %d of %d max_residual %.3e fp64_groups %d\n", s,
                     st.converged_columns, st.total_columns, st.max_residual, st.fp64_groups);
        write_all(dir + "/rho32cocg.bin", rho.data(), rho.size());
    }
    {   // iterations 0: construction check only; < 0: the reference's action timing loop
        gfa::green_solver_t solver(&plan, 0, 0, 64, 1);
        int const s0 = solver.solve(rho.data(), plan.nCols, 0);
        int const sb = solver.solve(rho.data(), plan.nCols, -3);
        std::fprintf(info, "zero_iterations status %d benchmark status %d flops %.6e\n", s0, sb,
                     double(plan.flops_performed));
    }
    for (void* p : {(void*)dx64, (void*)dy64, (void*)dx32, (void*)dy32, (void*)dcol, (void*)dbad}) HIPCHECK(hipFree(p));
    std::fclose(info);
    return 0;
}
