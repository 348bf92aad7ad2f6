// On-disk format of the tuned launch tables (the autotuners' results), kept
// free of HIP so it is testable on a CPU.
//
// One table per file, three lines:
//   clc-tuning-v2 <identity>        who wrote it: device architecture, CU
//                                   count, HIP runtime version
//   <v0> <v1> ... <vN-1>            the table
//   end <N> <fnv1a-64 of line 2>    a file cut short or edited is rejected
// A file whose identity differs from the reader's (another device, runtime
// or format) or that fails any check is ignored, and the caller re-tunes:
// the cache only ever saves time. Writers write a private temporary file
// and rename it over the target, so a reader sees the old table or the new
// one, never a mix. Same semantics as the reference's on-disk program cache
// (lib/src/Compute/ProgramCache.cpp:163-219: device-identity keys, stale or
// foreign files rebuilt, best-effort writes).
#pragma once

#include <string>
#include <vector>

namespace clc {

//! Reads the table at path; false when the file is missing, foreign (its
//! identity is not identity) or damaged. out is untouched on false.
bool read_tuning_file(const std::string& path, const std::string& identity,
                      std::vector<int>& out);
//! Writes values to path atomically; false on any failure (nothing is
//! thrown, and no temporary file is left behind).
bool write_tuning_file(const std::string& path, const std::string& identity,
                       const std::vector<int>& values);

} // namespace clc
