// cas-offinder: drop-in CLI over libcasoff.so with the reference's arguments
// (main.cpp:72-157):  cas-offinder {input|-} {C|G|A}[ids] {output|-} [--summary file]
#include <sys/time.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/casoff.h"

static void usage() {
    std::cout << "Cas-OFFinder 3.0.0 (MI355X / gfx950 HIP build)\n\n"
                 "Usage: cas-offinder [options] {input_filename|-} {C|G|A}[device_id(s)] {output_filename|-}\n"
                 "(G: MI355X GPU. C and A are not provided by this build.)\n\n"
                 "Options\n"
                 "  --summary <file>\t\t\t Print summary table to the specified file.\n\n"
                 "Example input file (DNA bulge 2, RNA bulge 1):\n"
                 "/var/chromosomes/human_grch38\n"
                 "NNNNNNNNNNNNNNNNNNNNNRG 2 1\n"
                 "GGCCGACCTGTCGCTGACGCNNN 5\n"
                 "CGCCAGCGTCAGCGACAGGTNNN 5\n"
                 "ACGGCGCCAGCGTCAGCGACNNN 5\n"
                 "GTCGCTGACGCTGGCGCCGTNNN 5\n";
}

static int die(const std::string &m) {
    std::cerr << m << std::endl;
    return 1;
}

static bool slurp(const char *path, std::string &out) {
    FILE *f = strcmp(path, "-") == 0 ? stdin : fopen(path, "rb");
    if (!f) return false;
    char buf[1 << 16];
    size_t r;
    while ((r = fread(buf, 1, sizeof buf, f)) > 0) out.append(buf, r);
    if (f != stdin) fclose(f);
    return true;
}

int main(int argc, char **argv) {
    struct timeval start, end;
    gettimeofday(&start, nullptr);
    std::string summary;
    std::vector<std::string> pos;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        if (a.size() > 2 && a[0] == '-' && a[1] == '-') {
            if (a.substr(2) == "summary" && i + 1 < argc) summary = argv[++i];
            else { usage(); return 0; }
        } else {
            pos.push_back(a);
        }
    }
    if (pos.size() < 3) { usage(); return 0; }
    const char dev = pos[1][0];
    if (dev != 'C' && dev != 'G' && dev != 'A') return die("Unknown option: " + pos[1]);
    if (dev != 'G') return die("Only G (the HIP GPU device) is provided by this build.");
    // device ids as the reference reads them (cas-offinder.cpp:131-153, :80-104);
    // this build searches on the lowest-numbered selected device
    int ndev = 0, bad = 0;
    cof_device_count(&ndev);
    std::vector<uint8_t> sel(ndev > 0 ? ndev : 1, 0);
    if (cof_device_ids(pos[1].c_str() + 1, ndev, sel.data(), &bad))
        return die("Something wrong with the device ID argument: " + pos[1]);
    for (int i = 0; i < bad; i++)
        std::cerr << "Something wrong with the device ID argument. Use all available devices instead..." << std::endl;
    int device = -1, nsel = 0;
    for (int d = 0; d < ndev; d++)
        if (sel[d]) {
            if (device < 0) device = d;
            nsel++;
        }
    if (device < 0) return die("No HIP devices found.");

    cof_ctx *ctx = nullptr;
    if (cof_open(device, &ctx)) return die(std::string("Device error: ") + cof_last_error(nullptr));
    std::cerr << "Total 1 device(s) found." << std::endl;
    if (nsel > 1) std::cerr << "Searching on device " << device << " (one device per process)." << std::endl;
    std::cerr << "Loading input file..." << std::endl;
    std::string input;
    if (!slurp(pos[0].c_str(), input)) return die("Critical error! Cannot read input file.");
    cof_query *q = nullptr;
    if (cof_query_parse(ctx, input.data(), input.size(), &q))
        return die(std::string("Critical error! ") + cof_last_error(ctx));
    const std::string &out = pos[2];
    {   // writeHeaders (cas-offinder.cpp:319-328)
        FILE *f = out == "-" ? stdout : fopen(out.c_str(), "wb");
        if (!f) return die("Cannot open output file " + out);
        fputs("##Generated by Cas-OFFinder 3.0.0\n"
              "#Id\tBulge Type\tcrRNA\tDNA\tChromosome\tLocation\tDirection\tMismatches\tBulge Size\n", f);
        if (f == stdout) fflush(f);
        else fclose(f);
    }
    const char *chrdir = nullptr;
    cof_query_info(q, &chrdir, nullptr, nullptr);
    int nfiles = 0;
    setenv("COF_VERBOSE", "1", 0);
    if (int rc = cof_genome_load(ctx, chrdir, &nfiles)) {
        std::cerr << cof_last_error(ctx) << std::endl;
        return rc == COF_E_IO ? 1 : 1;
    }
    if (nfiles > 0) {
        const char *m = getenv("COF_MODE");
        const int mode = (m && strcmp(m, "plain") == 0) ? COF_MODE_PLAIN : COF_MODE_BATCHED;
        std::cerr << "Searching the resident genome..." << std::endl;
        cof_result *r = nullptr;
        if (cof_search(ctx, &q, 1, mode, &r)) return die(std::string("Search failed: ") + cof_last_error(ctx));
        if (cof_result_format(ctx, r, 0, out.c_str(), 1, summary.c_str()))
            return die(std::string("Output failed: ") + cof_last_error(ctx));
        cof_result_free(r);
    }
    cof_query_free(q);
    cof_close(ctx);
    gettimeofday(&end, nullptr);
    const double secs = (end.tv_sec - start.tv_sec) + (end.tv_usec - start.tv_usec) * 1e-6;
    std::cerr << secs << " seconds elapsed." << std::endl;
    return 0;
}
