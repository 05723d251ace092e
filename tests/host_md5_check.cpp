// Test harness: MD5 of stdin with the product's host MD5 (host_md5.hpp),
// fed in pieces of the size given on the command line, printed as hex.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../par2rust_amd/csrc/host_md5.hpp"

int main(int argc, char **argv) {
    const size_t piece = argc > 1 ? (size_t)atol(argv[1]) : 4096;
    std::vector<unsigned char> buf(piece ? piece : 1);
    p2g::HostMd5 m;
    size_t got;
    while ((got = fread(buf.data(), 1, buf.size(), stdin)) > 0) m.update(buf.data(), got);
    unsigned char out[16];
    m.finish(out);
    for (int i = 0; i < 16; i++) printf("%02x", out[i]);
    printf("\n");
    return 0;
}
