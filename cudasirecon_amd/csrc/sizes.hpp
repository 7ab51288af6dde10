// Compile-time size lists of the specialised stage kernels.  A length outside
// these lists is rejected with a clear error at context creation.
#pragma once

// Lateral sizes of the input images (fused row kernels, overlap x pass).
#define SIMR_XY_SIZES(X)                                                                         \
  X(16) X(24) X(32) X(40) X(48) X(64) X(80) X(96) X(100) X(120) X(128) X(144) X(160) X(192)      \
  X(200) X(240) X(256) X(288) X(300) X(320) X(360) X(384) X(400) X(480) X(500) X(512) X(576)     \
  X(600) X(640) X(720) X(768) X(800) X(960) X(1000) X(1024) X(1152) X(1200) X(1280) X(1440) X(1536)   \
  X(1600) X(1920) X(2000) X(2048)

// Lateral sizes of the zoomed output grid (assembly x pass).
#define SIMR_ZOOM_SIZES(X)                                                                        \
  X(32) X(48) X(64) X(80) X(96) X(100) X(120) X(128) X(144) X(160) X(192) X(200) X(240) X(256)   \
  X(288) X(300) X(320) X(360) X(384) X(400) X(480) X(500) X(512) X(576) X(600) X(640) X(720)     \
  X(768) X(800) X(960) X(1000) X(1024) X(1152) X(1200) X(1280) X(1440) X(1536) X(1600) X(1920)   \
  X(2000) X(2048) X(2304) X(2400) X(2560) X(2880) X(3072) X(3200) X(3840) X(4000) X(4096)

// Axial sizes (z passes); 1 = 2D data.
#define SIMR_Z_SIZES(X)                                                                           \
  X(1) X(2) X(3) X(4) X(5) X(6) X(8) X(9) X(10) X(12) X(15) X(16) X(18) X(20) X(24) X(25) X(27)  \
  X(30) X(32) X(36) X(40) X(45) X(48) X(50) X(54) X(60) X(64) X(72) X(75) X(80) X(81) X(90)      \
  X(96) X(100) X(120) X(128) X(144) X(150) X(160) X(180) X(192) X(200) X(240) X(256) X(7) X(14)   \
  X(21) X(28) X(35) X(42) X(49) X(56) X(63) X(70) X(84) X(98) X(105) X(112) X(126) X(140) X(168)   \
  X(196) X(210) X(224)
