// wfa_gpu -- drop-in for the reference's `wfa_gpu` binary (wfa_gpu.cu:43-188).
//
//   wfa_gpu mismatch gap_opening gap_extension input_file wf_length check
//
// Same six positional arguments, same messages and exit codes, same stdout
// lines.  Alignment runs on the GPU through libtsunami_hip.so; check=1 adds
// the CPU verification pass of the reference (wfa_gpu.cu:150-185).
//
// Optional environment (the argument list stays identical to the reference):
//   TSUNAMI_MODE    plain (default) | batched | cigar
//   TSUNAMI_DEVICE  HIP device index (default 0)
//   TSUNAMI_JSON    path: also write a JSON summary there
//   TSUNAMI_CIGARS  path: write one CIGAR per line (mode cigar)
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../../include/tsunami_wfa.h"
#include "cpu_check.h"

namespace {

[[noreturn]] void die(const char* what, const char* detail) {
    // the reference prints "<error> in <file> at line <n>" and exits (wfa_gpu.cu:13-31)
    std::printf("%s: %s\n", what, detail ? detail : "");
    std::exit(EXIT_FAILURE);
}

const char* env_or(const char* name, const char* dflt) {
    const char* v = std::getenv(name);
    return (v && *v) ? v : dflt;
}

}  // namespace

int main(int argc, char const* argv[]) {
    if (argc != 7) {
        std::printf("ERROR! Please specify in order: mismatch, gap_opening, gap_extension, file name, wf_length, check\n");
        return 0;
    }
    const int mismatch = std::atoi(argv[1]);
    const int gap_opening = std::atoi(argv[2]);
    const int gap_extension = std::atoi(argv[3]);
    FILE* fp = std::fopen(argv[4], "r");
    const int wf_length = std::atoi(argv[5]);
    const int check = std::atoi(argv[6]);
    if (fp == nullptr) {
        std::printf("ERROR: Cannot open file.\n");
        return 1;
    }
    std::fclose(fp);

    const std::string mode = env_or("TSUNAMI_MODE", "plain");
    const int device = std::atoi(env_or("TSUNAMI_DEVICE", "0"));

    char err[512] = {0};
    tsu_dataset* ds = nullptr;
    int r = tsu_dataset_read(argv[4], &ds, err, sizeof(err));
    if (r != TSU_OK) die(tsu_strerror(r), err);
    const int64_t n = tsu_dataset_size(ds);

    tsu_ctx* ctx = nullptr;
    r = tsu_create(&ctx, device, mismatch, gap_opening, gap_extension);
    if (r != TSU_OK) die(tsu_strerror(r), "tsu_create (penalties must be mismatch>0, gap_opening>=0, gap_extension>0)");
    tsu_batch* b = nullptr;
    r = tsu_batch_from_dataset(ctx, ds, 0, n, &b);
    if (r != TSU_OK) die(tsu_strerror(r), tsu_last_error(ctx));

    std::vector<int32_t> score((size_t)n);
    tsu_timing tm;
    std::memset(&tm, 0, sizeof(tm));
    tsu_cigars* cg = nullptr;
    if (mode == "batched") r = tsu_align_batched(ctx, b, score.data(), &tm);
    else if (mode == "cigar") r = tsu_align_cigar(ctx, b, score.data(), &cg, &tm);
    else r = tsu_align_plain(ctx, b, wf_length, score.data(), &tm);
    if (r != TSU_OK) die(tsu_strerror(r), tsu_last_error(ctx));

    // GCUPS over equivalent cells p*t (wfa_gpu.cu:130-133), time = launch + sync
    const double gpu_s = tm.wall_ms / 1000.0;
    long double gcups = (long double)tm.cells;
    gcups /= 1E9L;
    gcups /= (long double)gpu_s;
    std::printf("GPU Time: %lf\n", gpu_s);
    std::printf("Estimated GCUPS GPU: : %Lf\n", gcups);

    if (const char* cpath = std::getenv("TSUNAMI_CIGARS")) {
        if (cg) {
            FILE* cf = std::fopen(cpath, "w");
            if (!cf) die("ERROR", "cannot write TSUNAMI_CIGARS file");
            std::vector<char> buf;
            for (int64_t i = 0; i < n; ++i) {
                const int need = tsu_cigars_string(cg, i, nullptr, 0);
                buf.resize((size_t)need + 1);
                tsu_cigars_string(cg, i, buf.data(), need + 1);
                std::fprintf(cf, "%d\t%s\n", score[(size_t)i], buf.data());
            }
            std::fclose(cf);
        }
    }

    bool equal = true;
    double cpu_s = 0;
    int threads = 0;
    if (check) {
        const auto t0 = std::chrono::high_resolution_clock::now();
        std::vector<int32_t> cpu;
        cpu_check_scores(ds, mismatch, gap_opening, gap_extension, cpu, &threads);
        const auto t1 = std::chrono::high_resolution_clock::now();
        cpu_s = std::chrono::duration<double>(t1 - t0).count();
        long double gcups_cpu = (long double)tm.cells;
        gcups_cpu /= 1E9L;
        gcups_cpu /= (long double)cpu_s;
        std::printf("CPU Time: %lf\n", cpu_s);
        std::printf("Estimated GCUPS SW: %Lf\n", gcups_cpu);
        for (int64_t i = 0; i < n && equal; ++i)
            if (cpu[(size_t)i] != score[(size_t)i]) equal = false;
        if (equal) std::printf("\n\n                    |ALIGNEMENT FINISHED CORRECTLY|\n");
        else std::printf("\n\n |ERROR|\n");
    }

    if (const char* jpath = std::getenv("TSUNAMI_JSON")) {
        FILE* jf = std::fopen(jpath, "w");
        if (jf) {
            char name[256] = {0};
            tsu_device_info(ctx, nullptr, nullptr, name, sizeof(name));
            std::fprintf(jf,
                         "{\"mode\": \"%s\", \"pairs\": %lld, \"gpu_s\": %.9f, \"kernel_ms\": %.6f, "
                         "\"alignments_per_s\": %.3f, \"gcups\": %.6f, \"check\": %d, \"cpu_s\": %.6f, "
                         "\"cpu_threads\": %d, \"equal\": %s, \"device\": \"%s\"}\n",
                         mode.c_str(), (long long)n, gpu_s, tm.kernel_ms, n / gpu_s, (double)gcups, check, cpu_s,
                         threads, equal ? "true" : "false", name);
            std::fclose(jf);
        }
    }
    tsu_cigars_free(cg);
    tsu_batch_free(b);
    tsu_destroy(ctx);
    tsu_dataset_free(ds);
    return 0;
}
