// crt_decrypt.cpp -- a C++ caller of include/xmp.h, linked against
// libxmp_amd.so: RSA-encrypt a batch of messages under one key held as
// one-instance arrays (read by every instance, i % 1), decrypt them with the
// Chinese remainder theorem -- c mod p, c^dp mod p, c mod q, c^dq mod q, then
// (mp*cp + mq*cq) mod n, the sequence of the reference's crt_rsa sample
// (samples/crt_rsa/rsa.cc:185-219) -- and check the result with
// xmpIntegersCmp and an export.  Exit status 0 when every message comes back.
#include <cstdint>
#include <cstdio>
#include <vector>

#include "xmp.h"

namespace {

struct Ints {
  xmpHandle_t h;
  xmpIntegers_t x = nullptr;
  Ints(xmpHandle_t handle, uint32_t bits, uint32_t count) : h(handle) {
    if (xmpIntegersCreate(h, &x, bits, count) != xmpErrorSuccess) x = nullptr;
  }
  ~Ints() {
    if (x) xmpIntegersDestroy(h, x);
  }
};

bool ok(xmpError_t e, const char *what) {
  if (e != xmpErrorSuccess) std::printf("%s: %s\n", what, xmpGetErrorString(e));
  return e == xmpErrorSuccess;
}

// one 64-bit value as two little-endian 32-bit limbs
void put64(std::vector<uint32_t> &v, uint64_t x) {
  v.push_back((uint32_t)x);
  v.push_back((uint32_t)(x >> 32));
}

}  // namespace

int main() {
  // the toy key of samples/crt_rsa/rsa.cc:167-178: n = 4931 * 3541
  const uint64_t n = 17460671, e = 65537, p = 4931, q = 3541, dp = 1063, dq = 113, cp = 15212136, cq = 2248536;
  const uint32_t count = 4096, bits = 64;
  xmpHandle_t h = nullptr;
  if (!ok(xmpHandleCreate(&h), "xmpHandleCreate")) return 1;
  int status = 1;
  {
    std::vector<uint32_t> msg;
    for (uint32_t i = 0; i < count; i++) put64(msg, (uint64_t)(i * 2654435761u) % n);
    Ints M(h, bits, count), C(h, bits, count), CM(h, bits, count), MP(h, bits, count), MQ(h, bits, count),
        SP(h, bits, count), SQ(h, bits, count), R(h, bits, count);
    Ints kn(h, bits, 1), ke(h, bits, 1), kp(h, bits, 1), kq(h, bits, 1), kdp(h, bits, 1), kdq(h, bits, 1),
        kcp(h, bits, 1), kcq(h, bits, 1);
    bool good = M.x && R.x && kcq.x;
    const uint64_t key[8] = {n, e, p, q, dp, dq, cp, cq};
    xmpIntegers_t kx[8] = {kn.x, ke.x, kp.x, kq.x, kdp.x, kdq.x, kcp.x, kcq.x};
    for (int k = 0; k < 8 && good; k++) {
      std::vector<uint32_t> v;
      put64(v, key[k]);
      good = ok(xmpIntegersImport(h, kx[k], 2, -1, sizeof(uint32_t), 0, 0, v.data(), 1), "import key");
    }
    good = good && ok(xmpIntegersImport(h, M.x, 2, -1, sizeof(uint32_t), 0, 0, msg.data(), count), "import");
    good = good && ok(xmpIntegersPowm(h, C.x, M.x, ke.x, kn.x, count), "encrypt");
    good = good && ok(xmpIntegersMod(h, CM.x, C.x, kp.x, count), "c mod p");
    good = good && ok(xmpIntegersPowm(h, MP.x, CM.x, kdp.x, kp.x, count), "mp");
    good = good && ok(xmpIntegersMod(h, CM.x, C.x, kq.x, count), "c mod q");
    good = good && ok(xmpIntegersPowm(h, MQ.x, CM.x, kdq.x, kq.x, count), "mq");
    good = good && ok(xmpIntegersMul(h, SP.x, MP.x, kcp.x, count), "mp*cp");
    good = good && ok(xmpIntegersMul(h, SQ.x, MQ.x, kcq.x, count), "mq*cq");
    good = good && ok(xmpIntegersAdd(h, SP.x, SP.x, SQ.x, count), "sum");
    good = good && ok(xmpIntegersMod(h, R.x, SP.x, kn.x, count), "mod n");
    std::vector<int32_t> cmp(count, 7);
    good = good && ok(xmpIntegersCmp(h, cmp.data(), R.x, M.x, count), "cmp");
    std::vector<uint32_t> back(2 * count, 0xFFFFFFFFu);
    uint32_t words = 0;
    good = good && ok(xmpIntegersExport(h, back.data(), &words, -1, sizeof(uint32_t), 0, 0, R.x, count), "export");
    uint32_t bad = 0;
    for (uint32_t i = 0; good && i < count; i++) bad += (cmp[i] != 0) || back[2 * i] != msg[2 * i] ||
                                                         back[2 * i + 1] != msg[2 * i + 1];
    if (good && words == 2 && bad == 0) {
      std::printf("crt ok: %u messages encrypted and decrypted\n", count);
      status = 0;
    } else if (good) {
      std::printf("crt FAILED: %u of %u messages differ (words %u)\n", bad, count, words);
    }
  }
  xmpHandleDestroy(h);
  return status;
}
