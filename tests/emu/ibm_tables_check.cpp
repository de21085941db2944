// Host check of the IBM table construction (cuda-lbm_amd/csrc/lbm_ibm_tables.hpp), built with
// AddressSanitizer and UndefinedBehaviorSanitizer: reads a marker set from a file, builds the
// tables and writes them out for the test to compare with its own construction.
//
// in:  int32 nx, ny, two_by_two, no_ds, n, has_ds; float32 xy[2n]; float32 ds[n] if has_ds
// out: int32 status (0 ok, 1 rejected), then on success int32 nb, n_ent, x0, x1, y0, y1,
//      band_node[nb], mk_ptr[n+1], mk_band[n_ent], float mk_w[n_ent], bn_ptr[nb+1],
//      bn_mk[n_ent], float bn_w[n_ent], box_map[(x1-x0)*(y1-y0)], kspread, mkf_band[16 n],
//      float mkf_w[16 n], bnf_mk[kspread nb], float bnf_w[kspread nb]
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../../cuda-lbm_amd/csrc/lbm_ibm_tables.hpp"

template <class T>
static void put(FILE* f, const std::vector<T>& v) {
    if (!v.empty()) std::fwrite(v.data(), sizeof(T), v.size(), f);
}

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]);
        return 2;
    }
    FILE* in = std::fopen(argv[1], "rb");
    if (!in) return 2;
    int32_t h[6];
    if (std::fread(h, 4, 6, in) != 6 || h[4] < 0) return 2;
    const int n = h[4];
    std::vector<float> xy(2 * size_t(n)), ds(h[5] ? size_t(n) : 0);
    if (std::fread(xy.data(), 4, xy.size(), in) != xy.size()) return 2;
    if (!ds.empty() && std::fread(ds.data(), 4, ds.size(), in) != ds.size()) return 2;
    std::fclose(in);
    lbm::IbmTables T;
    const char* why = lbm::ibm_build_tables(h[0], h[1], h[2] != 0, h[3] != 0, xy.data(), ds.empty() ? nullptr : ds.data(),
                                            n, T);
    FILE* out = std::fopen(argv[2], "wb");
    if (!out) return 2;
    int32_t status = why ? 1 : 0;
    std::fwrite(&status, 4, 1, out);
    if (!why) {
        int32_t hdr[6] = {int32_t(T.band_node.size()), int32_t(T.mk_band.size()), T.x0, T.x1, T.y0, T.y1};
        std::fwrite(hdr, 4, 6, out);
        put(out, T.band_node);
        put(out, T.mk_ptr);
        put(out, T.mk_band);
        put(out, T.mk_w);
        put(out, T.bn_ptr);
        put(out, T.bn_mk);
        put(out, T.bn_w);
        put(out, T.box_map);
        int32_t ks = T.kspread;
        std::fwrite(&ks, 4, 1, out);
        put(out, T.mkf_band);
        put(out, T.mkf_w);
        put(out, T.bnf_mk);
        put(out, T.bnf_w);
    } else {
        std::fprintf(stderr, "rejected: %s\n", why);
    }
    std::fclose(out);
    return 0;
}
