// hgx_evolve -- the paclet's process-isolated engine binary (paclet_source/hg_evolve_main.cpp)
// on this library, with the same command line and framing:
//
//   (no arguments)         one-shot: a WXF job on stdin, the WXF result on stdout; exit 0, or 1
//                          with the reason on stderr
//   --serve                worker: frames [u64 little-endian length][WXF job] on stdin until EOF,
//                          each answered by a frame on stdout
//   --serve-socket [FILE]  the same frames over one loopback TCP connection; the OS-assigned port
//                          is written to FILE (through FILE.tmp and a rename) or printed as
//                          "HGPORT <port>"
//   --device N             the GPU to use (default 0)
//
// A zero-length answer frame means that job failed (reason on stderr); a frame above 1 GiB is
// answered that way too. The device context lives as long as the worker, which is what the
// persistent worker is for: the paclet pays context creation once.
#include <arpa/inet.h>
#include <netinet/in.h>
#include <sys/socket.h>
#include <unistd.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "../../include/hgx.h"

namespace {

constexpr uint64_t kMaxJobBytes = 1ull << 30;

const char* const kUsage =
    "usage: hgx_evolve [--serve | --serve-socket [PORTFILE]] [--device N] [--version] [--help]\n";

std::string last_error(hgx_ctx* ctx) {
    char buf[1024] = {0};
    hgx_last_error(ctx, buf, sizeof(buf));
    return buf;
}

// One job: true and the result bytes, or false with the reason on stderr.
bool run(hgx_ctx* ctx, const std::vector<uint8_t>& job, std::vector<uint8_t>& result) {
    uint8_t* out = nullptr;
    size_t n = 0;
    if (hgx_run_wxf(ctx, job.data(), job.size(), &out, &n) != HGX_OK) {
        std::fprintf(stderr, "hgx_evolve: job failed: %s\n", last_error(ctx).c_str());
        std::fflush(stderr);
        return false;
    }
    result.assign(out, out + n);
    hgx_buffer_free(out);
    return true;
}

using ReadFn = std::function<bool(uint8_t*, size_t)>;
using WriteFn = std::function<bool(const uint8_t*, size_t)>;

// The frame loop shared by --serve and --serve-socket; ends at end of input.
int serve(hgx_ctx* ctx, const ReadFn& rd, const WriteFn& wr) {
    std::vector<uint8_t> job, result;
    for (;;) {
        uint8_t hdr[8];
        if (!rd(hdr, 8)) return 0;                               // end of input between frames
        uint64_t n = 0;
        for (int i = 0; i < 8; ++i) n |= (uint64_t)hdr[i] << (8 * i);
        bool ok = false;
        if (n > kMaxJobBytes) {
            std::fprintf(stderr, "hgx_evolve: job frame of %llu bytes exceeds the 1 GiB cap\n",
                         (unsigned long long)n);
            result.clear();
        } else {
            job.resize((size_t)n);
            if (n && !rd(job.data(), (size_t)n)) {
                std::fprintf(stderr, "hgx_evolve: truncated job frame\n");
                return 1;
            }
            ok = run(ctx, job, result);
        }
        const uint64_t m = ok ? result.size() : 0;
        uint8_t out[8];
        for (int i = 0; i < 8; ++i) out[i] = (uint8_t)(m >> (8 * i));
        if (!wr(out, 8) || (m && !wr(result.data(), (size_t)m))) return 1;
    }
}

int serve_socket(hgx_ctx* ctx, const char* portfile) {
    const int ls = socket(AF_INET, SOCK_STREAM, 0);
    if (ls < 0) { std::fprintf(stderr, "hgx_evolve: socket() failed\n"); return 1; }
    sockaddr_in addr{};
    addr.sin_family = AF_INET;
    addr.sin_addr.s_addr = htonl(INADDR_LOOPBACK);
    addr.sin_port = 0;                                          // an OS-assigned free port
    socklen_t alen = sizeof(addr);
    if (bind(ls, (sockaddr*)&addr, sizeof(addr)) != 0 || getsockname(ls, (sockaddr*)&addr, &alen) != 0 ||
        listen(ls, 1) != 0) {
        std::fprintf(stderr, "hgx_evolve: cannot listen on the loopback interface\n");
        close(ls);
        return 1;
    }
    const unsigned port = ntohs(addr.sin_port);
    if (portfile) {
        const std::string tmp = std::string(portfile) + ".tmp";
        FILE* f = std::fopen(tmp.c_str(), "w");
        if (!f || std::fprintf(f, "%u\n", port) < 0 || std::fclose(f) != 0 ||
            std::rename(tmp.c_str(), portfile) != 0) {
            std::fprintf(stderr, "hgx_evolve: cannot write the port file\n");
            close(ls);
            return 1;
        }
    } else {
        std::printf("HGPORT %u\n", port);
        std::fflush(stdout);
    }
    const int cs = accept(ls, nullptr, nullptr);
    if (cs < 0) { std::fprintf(stderr, "hgx_evolve: accept() failed\n"); close(ls); return 1; }
    auto rd = [cs](uint8_t* p, size_t n) {
        while (n) {
            const ssize_t r = recv(cs, p, n, 0);
            if (r <= 0) return false;
            p += r; n -= (size_t)r;
        }
        return true;
    };
    auto wr = [cs](const uint8_t* p, size_t n) {
        while (n) {
            const ssize_t r = send(cs, p, n, MSG_NOSIGNAL);
            if (r <= 0) return false;
            p += r; n -= (size_t)r;
        }
        return true;
    };
    const int rc = serve(ctx, rd, wr);
    close(cs);
    close(ls);
    return rc;
}

}  // namespace

int main(int argc, char** argv) {
    bool frames = false, sock = false;
    const char* portfile = nullptr;
    int device = 0;
    for (int i = 1; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--serve")) frames = true;
        else if (!std::strcmp(argv[i], "--serve-socket")) {
            sock = true;
            if (i + 1 < argc && argv[i + 1][0] != '-') portfile = argv[++i];
        } else if (!std::strcmp(argv[i], "--device") && i + 1 < argc) device = std::atoi(argv[++i]);
        else if (!std::strcmp(argv[i], "--version")) { std::puts("hgx_evolve (hgx multiway evolution, gfx950)"); return 0; }
        else if (!std::strcmp(argv[i], "--help") || !std::strcmp(argv[i], "-h")) { std::fputs(kUsage, stdout); return 0; }
        else {
            std::fprintf(stderr, "hgx_evolve: unknown argument '%s'\n%s", argv[i], kUsage);
            return 2;
        }
    }
    hgx_ctx* ctx = nullptr;
    if (hgx_ctx_create(device, 0, &ctx) != HGX_OK) {
        std::fprintf(stderr, "hgx_evolve: %s\n", last_error(nullptr).c_str());
        return 1;
    }
    int status = 0;
    if (sock) {
        status = serve_socket(ctx, portfile);
    } else if (frames) {
        auto rd = [](uint8_t* p, size_t n) { return std::fread(p, 1, n, stdin) == n; };
        auto wr = [](const uint8_t* p, size_t n) {
            const bool ok = std::fwrite(p, 1, n, stdout) == n;
            std::fflush(stdout);
            return ok;
        };
        status = serve(ctx, rd, wr);
    } else {
        std::vector<uint8_t> job, result;
        uint8_t buf[65536];
        size_t r;
        while ((r = std::fread(buf, 1, sizeof(buf), stdin)) > 0) job.insert(job.end(), buf, buf + r);
        if (run(ctx, job, result)) {
            std::fwrite(result.data(), 1, result.size(), stdout);
            std::fflush(stdout);
        } else {
            status = 1;
        }
    }
    hgx_ctx_destroy(ctx);
    return status;
}
