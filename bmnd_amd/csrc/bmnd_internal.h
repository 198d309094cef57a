// Shared by the library's translation units (not part of the C ABI).
#pragma once

// records msg as the calling thread's last error (bmnd_hip_last_error) and
// returns code
__attribute__((visibility("hidden"))) int bmnd_internal_fail(int code, const char *msg);
