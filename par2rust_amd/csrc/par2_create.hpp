// par2_create.hpp -- C++ host restatement of par2rust's create API on top of
// the GPU engine: CreateOptions / VolumeScheme (src/creator.rs:110-173),
// SourceFile (src/source.rs:27-47), SourceFile::scan (src/source.rs:53-136)
// and run_create (src/creator.rs:182-321).  Errors are p2g status codes,
// which map one-to-one onto Par2Error (src/error.rs:7-41).
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/par2gpu.h"

namespace par2 {

using Md5Hash = std::array<uint8_t, 16>;

struct SliceChecksum {
    uint8_t md5[16];
    uint32_t crc32;
};

struct SourceFile {
    std::vector<uint8_t> name;  // display name as stored in the packet
    std::string path;
    uint64_t length = 0;
    uint8_t hash_full[16] = {};
    uint8_t hash16k[16] = {};
    uint8_t file_id[16] = {};
    std::vector<SliceChecksum> slice_checksums;
};

struct VolumeScheme {
    int kind = P2G_SCHEME_SINGLE;  // library default is Single, src/creator.rs:163-173
    uint32_t uniform_count = 0;
    // Limited { max_blocks_per_volume, inner } wrappers around the scheme
    // (src/creator.rs:131-134), their caps from the innermost outward
    std::vector<uint32_t> caps;
    std::vector<uint32_t> explicit_sizes;
};

struct CreateOptions {
    std::string output;
    uint64_t slice_size = 4096;
    uint32_t recovery_block_count = 0;
    VolumeScheme volume_scheme;
};

struct Error {
    int code = P2G_OK;
    std::string message;
};

// SourceFile::scan: per-slice MD5+CRC-32 on the GPU, hash_full / hash16k on
// the host while the GPU works.
int scan(p2g_ctx *ctx, const std::string &path, const std::vector<uint8_t> &name,
         uint64_t slice_size, SourceFile &out, Error &err);

// run_create: index file, then one encode over every source and the volume
// files.  `written` receives the paths written, index first.
int run_create(p2g_ctx *ctx, const CreateOptions &opts, const std::vector<SourceFile> &sources,
               std::vector<std::string> &written, Error &err);

// scan + run_create with every file read once into HBM and one fused device
// pass (the fused shape of src/creator.rs:447-700).  *declined: the set does
// not fit the free HBM; nothing was written and the caller streams instead.
int create_resident(p2g_ctx *ctx, const CreateOptions &opts, const std::vector<std::string> &paths,
                    const std::vector<std::vector<uint8_t>> &names, std::vector<std::string> &written,
                    Error &err, bool *declined);

// resolve_volume_sizes, src/creator.rs:325-431 (par2_scheme.cpp).
int resolve_volume_sizes(const VolumeScheme &s, uint32_t total, std::vector<uint32_t> &out,
                         Error &err);

// derive_volume_filename, src/creator.rs:1046-1057 (par2_scheme.cpp).
std::string volume_name(const std::string &index, uint32_t first, uint32_t count);

}  // namespace par2
