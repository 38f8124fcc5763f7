// Grid sizes n that libdns_hip carries kernels for: every even n = 2^a 3^b
// from 8 to 1024.  X(n) is expanded once per size: dns_api.hip builds its
// dispatch from this list, and the build compiles inst.hip once per size
// (-DDNS_N=<n>), so the sizes compile in parallel.
#pragma once
#define DNS_GRID_SIZES(X)                                                                                  \
  X(8) X(12) X(16) X(18) X(24) X(32) X(36) X(48) X(54) X(64) X(72) X(96) X(108) X(128) X(144) X(162)      \
  X(192) X(216) X(256) X(288) X(324) X(384) X(432) X(486) X(512) X(576) X(648) X(768) X(864) X(972)     \
  X(1024)
