// generate_dataset -- writes ./sequences.txt like the reference's generator
// (generate_dataset.cpp:11-75), but seeded.
//
//   generate_dataset num_couples sequence_length error_rate [seed] [mode]
//
// With the reference's three arguments the pairs follow its model: equal
// lengths, substitutions only, length * error_rate / 100 (rounded down) of
// them per pair at distinct positions, each to a different base
// (generate_dataset.cpp:20-23,45-66), so the reference's run scripts and
// their wf_length values apply unchanged.  mode "mixed" makes each edit a
// substitution, an insertion or a deletion with probability 1/3 and takes
// round(error_rate / 100 * length) edits (the build's benchmark workloads);
// mode "sub" is the default.  error_rate is a percentage as in the
// reference.  Seed defaults to 0x5453554E.  Set TSUNAMI_LEN_MAX /
// TSUNAMI_ERR_MAX to draw lengths log-uniformly in [length, LEN_MAX] and
// rates uniformly in [error_rate, ERR_MAX] percent.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

#include "../../include/tsunami_wfa.h"

int main(int argc, char* argv[]) {
    if (argc < 4) {
        std::cout << "Required arguments: number of couples, sequences length, error rate" << std::endl;
        std::exit(1);
    }
    const long long num = std::atoll(argv[1]);
    const int len = std::atoi(argv[2]);
    const double rate = std::atof(argv[3]);
    tsu_gen_cfg cfg;
    std::memset(&cfg, 0, sizeof(cfg));
    cfg.seed = argc > 4 ? std::strtoull(argv[4], nullptr, 0) : 0x5453554Eull;
    cfg.len_min = len;
    cfg.len_max = len;
    cfg.err_min = rate / 100.0;
    cfg.err_max = rate / 100.0;
    cfg.edit_mode = 2;
    if (argc > 5) {
        const std::string mode = argv[5];
        if (mode == "mixed") {
            cfg.edit_mode = 0;
        } else if (mode != "sub") {
            std::cout << "generate_dataset: mode must be \"sub\" or \"mixed\"" << std::endl;
            return 1;
        }
    }
    if (const char* v = std::getenv("TSUNAMI_LEN_MAX")) cfg.len_max = std::atoi(v);
    if (const char* v = std::getenv("TSUNAMI_ERR_MAX")) cfg.err_max = std::atof(v) / 100.0;
    tsu_dataset* ds = nullptr;
    int r = tsu_dataset_generate(&cfg, num, &ds);
    if (r != TSU_OK) {
        std::cout << "generate_dataset: " << tsu_strerror(r) << std::endl;
        return 1;
    }
    r = tsu_dataset_write(ds, "sequences.txt", (int)(rate + 0.5));
    tsu_dataset_free(ds);
    if (r != TSU_OK) {
        std::cout << "generate_dataset: cannot write sequences.txt" << std::endl;
        return 1;
    }
    return 0;
}
