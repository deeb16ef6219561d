// wave_fft.h -- one 64-lane wave transforms one line of n = 64 E complex points.
//
// Lane l holds entries l + 64 m (m = 0 .. E-1) of the line before and after
// the transform ("distribution D"), so loads, stores and element-wise work
// around the transform need no data exchange.  The transform is a Stockham
// autosort FFT of three passes:
//   pass 1  radix E, stride 1      registers -> LDS
//   pass 2  radix 8, stride E      LDS -> LDS   (E / 8 butterflies per lane)
//   pass 3  radix 8, stride n / 8  LDS -> registers, in distribution D
// Each pass is split into a read phase (every lane loads all its inputs) and
// a write phase, separated by a barrier, so one LDS line buffer suffices.  LDS
// indices are padded by one entry every E (i + i / E): the stride-E writes of
// pass 1 then fall on distinct banks.  Twiddles come from the FP64-built table
// W_Ntab^m of the system (read through the L1); the radix-8/16/32 butterflies
// use exact constant roots.  Sign: forward e^{-i}, inverse e^{+i} (FFTW).
#pragma once

#if defined(__HIPCC__)
#define CH_HD __host__ __device__ __forceinline__
#else
#define CH_HD inline
#endif

namespace chd {

// cos(pi k / 16), k = 0 .. 8 (W_32 roots)
CH_HD double cos_pi16(int k) {
    switch(k) {
    case 0: return 1.0;
    case 1: return 0.98078528040323044912618223613424;
    case 2: return 0.92387953251128675612818318939679;
    case 3: return 0.83146961230254523707878837761791;
    case 4: return 0.70710678118654752440084436210485;
    case 5: return 0.55557023301960222474283081394853;
    case 6: return 0.38268343236508977172845998403040;
    case 7: return 0.19509032201612826784828486847702;
    default: return 0.0;
    }
}

// W_32^k = e^{-2 pi i k / 32} (forward) as (cos, -sin); exact for k = 0 .. 31
template <typename T2>
CH_HD T2 root32(int k, bool inv) {
    using T = decltype(T2{}.x);
    k &= 31;
    const int q = k >> 3, r = k & 7;  // angle = (q pi/2) + r pi/16
    double c = cos_pi16(r), s = cos_pi16(8 - r);  // cos, sin of r pi/16
    double cc, ss;
    switch(q) {
    case 0: cc = c; ss = s; break;
    case 1: cc = -s; ss = c; break;
    case 2: cc = -c; ss = -s; break;
    default: cc = s; ss = -c; break;
    }
    return T2{(T)cc, (T)(inv ? ss : -ss)};
}

template <typename T2>
CH_HD T2 w_mul(T2 a, T2 b) {
    return T2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// in-place DFT of size R (2, 4, 8, 16, 32) on v[0 .. R-1], radix-2 DIT recursion
template <int R, bool INV, typename T2>
CH_HD void dft(T2 *v) {
    if constexpr(R == 1) {
        return;
    }
    else if constexpr(R == 2) {
        const T2 a = v[0], b = v[1];
        v[0] = T2{a.x + b.x, a.y + b.y};
        v[1] = T2{a.x - b.x, a.y - b.y};
    }
    else if constexpr(R == 4) {
        const T2 s02{v[0].x + v[2].x, v[0].y + v[2].y}, d02{v[0].x - v[2].x, v[0].y - v[2].y};
        const T2 s13{v[1].x + v[3].x, v[1].y + v[3].y}, d13{v[1].x - v[3].x, v[1].y - v[3].y};
        // forward: (-i) d13; inverse: (+i) d13
        const T2 t = INV ? T2{-d13.y, d13.x} : T2{d13.y, -d13.x};
        v[0] = T2{s02.x + s13.x, s02.y + s13.y};
        v[2] = T2{s02.x - s13.x, s02.y - s13.y};
        v[1] = T2{d02.x + t.x, d02.y + t.y};
        v[3] = T2{d02.x - t.x, d02.y - t.y};
    }
    else {
        constexpr int H = R / 2;
        T2 e[H], o[H];
#pragma unroll
        for(int i = 0; i < H; i++) {
            e[i] = v[2 * i];
            o[i] = v[2 * i + 1];
        }
        dft<H, INV>(e);
        dft<H, INV>(o);
#pragma unroll
        for(int k = 0; k < H; k++) {
            const T2 t = k == 0 ? o[k] : w_mul(o[k], root32<T2>(k * (32 / R), INV));
            v[k] = T2{e[k].x + t.x, e[k].y + t.y};
            v[k + H] = T2{e[k].x - t.x, e[k].y - t.y};
        }
    }
}

template <typename T2, int LOG2N>
struct WaveFFT {
    static constexpr int n = 1 << LOG2N;
    static constexpr int E = n / 64;   // entries per lane
    static constexpr int B = E / 8;    // radix-8 butterflies per lane in passes 2, 3
    static constexpr int LDS = n + n / E;
    static_assert(E == 8 || E == 16 || E == 32, "wave FFT lines hold 512, 1024 or 2048 points");

    CH_HD static int pad(int i) { return i + i / E; }

    // table[m] = W_Ntab^m; tstride = Ntab / n
    template <bool INV>
    CH_HD static T2 tw(const T2 *table, int m, int tstride) {
        const T2 w = table[m * tstride];
        return INV ? T2{w.x, -w.y} : w;
    }

    // pass 1: v in distribution D (input index lane + 64 r); outputs lane E + r
    template <bool INV>
    CH_HD static void pass1_write(int lane, T2 (&v)[E], T2 *lds) {
        dft<E, INV>(v);
#pragma unroll
        for(int r = 0; r < E; r++) lds[pad(lane * E + r)] = v[r];
    }
    // passes 2 and 3 read the inputs j + r n / 8 of butterflies j = lane + 64 b
    CH_HD static void pass_read(int lane, T2 (&v)[E], const T2 *lds) {
#pragma unroll
        for(int b = 0; b < B; b++)
#pragma unroll
            for(int r = 0; r < 8; r++) v[b * 8 + r] = lds[pad(lane + 64 * b + r * (n / 8))];
    }
    // pass 2: radix 8, stride Ns = E; twiddle W_{8E}^{jm r} = W_n^{8 jm r}
    template <bool INV>
    CH_HD static void pass2_write(int lane, T2 (&v)[E], T2 *lds, const T2 *table, int tstride) {
#pragma unroll
        for(int b = 0; b < B; b++) {
            const int j = lane + 64 * b, jm = j % E;
#pragma unroll
            for(int r = 1; r < 8; r++) v[b * 8 + r] = w_mul(v[b * 8 + r], tw<INV>(table, 8 * jm * r, tstride));
            dft<8, INV>(&v[b * 8]);
#pragma unroll
            for(int r = 0; r < 8; r++) lds[pad((j - jm) * 8 + jm + r * E)] = v[b * 8 + r];
        }
    }
    // pass 3: radix 8, stride n / 8 (jm = j); twiddle W_n^{j r}; the outputs
    // j + r n / 8 = lane + 64 (b + r B) land in distribution D
    template <bool INV>
    CH_HD static void pass3_final(int lane, T2 (&v)[E], const T2 *table, int tstride) {
        T2 y[E];
#pragma unroll
        for(int b = 0; b < B; b++) {
            const int j = lane + 64 * b;
#pragma unroll
            for(int r = 1; r < 8; r++) v[b * 8 + r] = w_mul(v[b * 8 + r], tw<INV>(table, j * r, tstride));
            dft<8, INV>(&v[b * 8]);
#pragma unroll
            for(int r = 0; r < 8; r++) y[b + r * B] = v[b * 8 + r];
        }
#pragma unroll
        for(int m = 0; m < E; m++) v[m] = y[m];
    }
};

// 256-point transform on one wave, 4 entries per lane (distribution D: lane
// l holds entries l + 64 m), as four radix-4 Stockham passes: pass s (stride
// Ns = 4^s) reads inputs j + 64 r of butterfly j = lane, applies W_{4 Ns}^{jm r}
// (jm = j mod Ns) and writes outputs (j - jm) 4 + jm + r Ns; the last pass's
// outputs lane + 64 r are distribution D again.  The wave's own LDS buffer
// (LDS entries, padded by one every four) carries the three exchanges; the
// caller orders the block's earlier accesses to it.  tstride: table index of
// W_256^1 in the table of W_Ntab^m.
template <typename T2>
struct WaveFFT4 {
    static constexpr int n = 256, E = 4, LDS = 320;
    CH_HD static int pad(int i) { return i + (i >> 2); }
    template <bool INV>
    CH_HD static T2 tw(const T2 *table, int m, int tstride) {
        const T2 w = table[m * tstride];
        return INV ? T2{w.x, -w.y} : w;
    }
    template <bool INV>
    __device__ static void run(int lane, T2 (&v)[4], T2 *lds, const T2 *table, int tstride) {
        dft<4, INV>(v);
#pragma unroll
        for(int r = 0; r < 4; r++) lds[pad(4 * lane + r)] = v[r];
#pragma unroll
        for(int s = 1; s < 4; s++) {
            const int Ns = 1 << (2 * s);
            wave_lds_sync();
#pragma unroll
            for(int r = 0; r < 4; r++) v[r] = lds[pad(lane + 64 * r)];
            const int jm = lane & (Ns - 1);
#pragma unroll
            for(int r = 1; r < 4; r++) v[r] = w_mul(v[r], tw<INV>(table, jm * r * (64 >> (2 * s)), tstride));
            dft<4, INV>(v);
            if(s < 3) {
                wave_lds_sync();
#pragma unroll
                for(int r = 0; r < 4; r++) lds[pad((lane - jm) * 4 + jm + r * Ns)] = v[r];
            }
        }
    }
};

}  // namespace chd
