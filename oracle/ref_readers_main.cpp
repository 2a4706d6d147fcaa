// Driver for the reference's own genome readers (read_fasta.cpp, read_twobit.cpp
// compiled from the reference tree).  Test infrastructure: it prints what the
// reference's run_cas_offinder (main.cpp:49-54) would hold for one file, so the
// oracle's and the product's loaders can be checked against it.
// Output: "status <0|1|2>" then, for status 0, "nchr <n>", one "chr <pos> <name>"
// line per record and "len <bytes>" followed by the raw text on the next line.
#include <cstdio>
#include <string>
#include <vector>
#include "read_fasta.h"
#include "read_twobit.h"

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    std::string path = argv[1];
    std::vector<std::string> names;
    std::string content;
    std::vector<unsigned long long> pos;
    int status = 0;
    if (read_fasta(path, names, content, pos)) {
        if (read_twobit(path, names, content, pos)) status = 1;
    }
    printf("status %d\n", status);
    if (status) return 0;
    printf("nchr %zu\n", names.size());
    for (size_t i = 0; i < names.size(); i++)
        printf("chr %llu %s\n", i < pos.size() ? pos[i] : 0ull, names[i].c_str());
    printf("len %zu\n", content.size());
    fwrite(content.data(), 1, content.size(), stdout);
    printf("\n");
    return 0;
}
