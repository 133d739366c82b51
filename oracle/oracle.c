/*
 * oracle.c -- sequential CPU restatement of goetia's dBG hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker for the HIP product
 * path.  Never linked into, or called by, the product library.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* The 256 random words of the vendored cyclic hash,
 * include/goetia/hashing/rollinghash/characterhash.h:27-113 (same values,
 * kept here in decimal so the oracle does not share the product's copy). */
static const uint64_t OR_T[256] = {
    2821057300106699797ULL, 11325163037179593313ULL, 5416224960711267644ULL,
    7528795173630926414ULL, 13801220329386839123ULL, 9571312227480569745ULL,
    13287641504033217005ULL, 17026797478042468886ULL, 12854026315321965989ULL,
    6374470260133235413ULL, 2204210189948696549ULL, 16294121409648186942ULL,
    52948147103884203ULL, 6415147802595498673ULL, 16198407629368318724ULL,
    11766006283457710847ULL, 5173278555550225963ULL, 291850381695702383ULL,
    9067924282960618557ULL, 7697914929610629098ULL, 12522305300226911288ULL,
    9498027593538051549ULL, 17257649845518499103ULL, 14516692441855410225ULL,
    8715626173811934569ULL, 5325903972055096315ULL, 16166526287333920336ULL,
    2569195063296827853ULL, 17109513377892528534ULL, 13886494041707436679ULL,
    14893740505833355047ULL, 4234952225844646826ULL, 14358747021030577240ULL,
    1887737051715844916ULL, 2330098280227001344ULL, 9188490525853320411ULL,
    6025534915685878506ULL, 17424697768797008296ULL, 7637984810305105866ULL,
    12367942525066656103ULL, 9840387827592398565ULL, 10533248810674727780ULL,
    13207324779243726126ULL, 7315623160354210225ULL, 11948867329934421237ULL,
    3644170777857093101ULL, 9676205714709334695ULL, 14771130086372921028ULL,
    2292394639782612013ULL, 10197765460546535208ULL, 14073391231741944415ULL,
    8621873250296467544ULL, 6590463839587788851ULL, 13850038707066944250ULL,
    18429044708838245451ULL, 215258643983951745ULL, 1884062495295395303ULL,
    12663706866596696824ULL, 14449601608250063354ULL, 3435894449503459935ULL,
    10870768431824442842ULL, 11504721573538595258ULL, 12186578141009955262ULL,
    16198870037728999207ULL, 10669405064869886793ULL, 16664410744025174816ULL,
    18397359403060890687ULL, 15956807086001210932ULL, 14617266033040342736ULL,
    13852442956164108892ULL, 10306008263506153018ULL, 9404339731978646439ULL,
    8655922979817266314ULL, 12809422002881587418ULL, 4261482358731270443ULL,
    7732774012761317500ULL, 16502580177155673326ULL, 15611680917896763796ULL,
    10780076607752328646ULL, 1542572181126078483ULL, 3654327363130598725ULL,
    16794808258298155332ULL, 12369059170961750015ULL, 3132818342866713807ULL,
    836480985777824379ULL, 13906561201785132592ULL, 9858196578848854067ULL,
    4372375965812152942ULL, 14254521215283273274ULL, 7596145618544622863ULL,
    16342747097867571709ULL, 1980540366340112804ULL, 8262177221342635774ULL,
    13125315094409684063ULL, 17203349866499262586ULL, 16759515313626684424ULL,
    15342020348994741163ULL, 2706664420015103890ULL, 11384314333735679458ULL,
    15295060863465421772ULL, 6664089646357471419ULL, 14233415411242489719ULL,
    12908263413292462998ULL, 1353330147242825292ULL, 15303480952098533508ULL,
    337359223552501433ULL, 1568788723399388210ULL, 17644968499059613652ULL,
    3847406208385860132ULL, 17523265639259566024ULL, 12238335531163888847ULL,
    3903849605456502145ULL, 8263775787236913822ULL, 12389831137335901677ULL,
    1292775400312032391ULL, 3771478312966411459ULL, 4541792594786962408ULL,
    14808262216606926175ULL, 2581919202420840491ULL, 166168189331466451ULL,
    6554319341342777275ULL, 17145090797639572225ULL, 12732027427133868456ULL,
    14559633844336852506ULL, 909502806480392211ULL, 5778876322733837940ULL,
    16859623521438723622ULL, 5134076461472812216ULL, 12354633801538016303ULL,
    17850322572567648614ULL, 5820236151567115665ULL, 6576849339431981618ULL,
    10590600560778871269ULL, 10940127808752622216ULL, 8565035553835162937ULL,
    1774065049477478434ULL, 4814212482303396695ULL, 4275189062710062116ULL,
    630349411251672937ULL, 6436038361922976831ULL, 16664238356166903433ULL,
    2398603741284037446ULL, 11468261837055649518ULL, 17514985377544487208ULL,
    3177974761867308872ULL, 7692698080149068968ULL, 13837720152036376942ULL,
    5506266379279335520ULL, 5308518106289494774ULL, 16526183736085529801ULL,
    357257327638821318ULL, 1181833586544916844ULL, 16664823884918697019ULL,
    2529708085974818140ULL, 18394812950618671631ULL, 9885543901797367607ULL,
    10673912057166833585ULL, 522162561327397618ULL, 8433940106549935649ULL,
    2621653722353075094ULL, 9722505992377914047ULL, 1711872796729989241ULL,
    17798157461229462354ULL, 16913106389905575378ULL, 7896888628990885297ULL,
    15864715499398966611ULL, 14079164878488738891ULL, 3050246499307582612ULL,
    1703346441458602658ULL, 7310689803047170057ULL, 11498964545764371542ULL,
    17202646980119590472ULL, 9680839911901944918ULL, 7640258924910894043ULL,
    6529023846032485156ULL, 483507376575727765ULL, 13079525949120166236ULL,
    10170150048283242299ULL, 2109960879517206ULL, 16760628739394415400ULL,
    10582518701122574096ULL, 4355072925142789383ULL, 5781655723244612819ULL,
    7768064507395797426ULL, 921478671905296912ULL, 3179121294549908816ULL,
    9975781372392684305ULL, 549738065842820422ULL, 12640068765622852389ULL,
    8229728960599470880ULL, 9375512254211572010ULL, 12744798652893158085ULL,
    2363478901711075627ULL, 2116689541286966759ULL, 8180262042010236006ULL,
    4249443692261027106ULL, 2707168392994699624ULL, 5514287253458790481ULL,
    1814178923525761966ULL, 5577051504293005778ULL, 12771098131382550559ULL,
    12032790465375591411ULL, 2535960658173991567ULL, 10820030687998681802ULL,
    4979290776778324590ULL, 16457002048535804954ULL, 12244850006062512789ULL,
    7609998496390090313ULL, 14104243468864345707ULL, 720433609917385434ULL,
    13392299073419520539ULL, 4097474698105279986ULL, 8299413979248706606ULL,
    16290789562718246821ULL, 7335842782667280191ULL, 16257887162134443769ULL,
    17861273412307426758ULL, 13701592676399062223ULL, 17373825138829494763ULL,
    7944986297238901628ULL, 9835163643181430505ULL, 366723710453002021ULL,
    15396126198530254709ULL, 11511321399930066684ULL, 3540964612902366865ULL,
    5036883055268744511ULL, 4962193520082228875ULL, 14120707866588336117ULL,
    18241267409277997530ULL, 17669943049921343994ULL, 9932182549666343663ULL,
    9558227449175976386ULL, 1092844097255405105ULL, 1904183829110741896ULL,
    7155690151231400523ULL, 10439234361950944381ULL, 1361222343450861792ULL,
    9095344136183365091ULL, 4898471236227651240ULL, 2589681900905316607ULL,
    12762880031044612488ULL, 9939678776623939418ULL, 505211974413939563ULL,
    9972319229645889248ULL, 7319067952382618431ULL, 8681790484606533729ULL,
    6965201783339143147ULL, 7006134092273617374ULL, 5342723917871999735ULL,
    984365199659547641ULL, 5559120030223996309ULL, 9939414356931300058ULL,
    1217488313221569951ULL, 17323850086388522307ULL, 15532728393800726547ULL,
    13935833578031356880ULL,
};

uint64_t or_table_value(int c) { return OR_T[(unsigned char)c]; }

/* cyclichash.h:41-43 fastleftshiftn: ((x & maskn) << r) | (x >> (64-r)).
 * For r == 0 the reference shifts by 64, which x86-64 masks to 0, giving x. */
uint64_t or_rotl(uint64_t x, unsigned r) {
    r &= 63u;
    return r ? ((x << r) | (x >> (64u - r))) : x;
}

uint64_t or_rotr(uint64_t x, unsigned r) {
    r &= 63u;
    return r ? ((x >> r) | (x << (64u - r))) : x;
}

/* alphabets.hh:132-145 DNA_SIMPLE::_complement */
char or_complement(char c) {
    switch (c) {
        case 'A': return 'T';
        case 'T': return 'A';
        case 'C': return 'G';
        case 'G': return 'C';
        default:  return '\0';
    }
}

/* One CyclicHash<uint64_t,unsigned char> instance: cyclichash.h:23-138. */
typedef struct { uint64_t h; unsigned myr; } or_cyclic;

static void cyc_reset(or_cyclic* c) { c->h = 0; }                       /* :124-126 */
static void cyc_eat(or_cyclic* c, char in) {                             /* :105-108 */
    c->h = or_rotl(c->h, 1) ^ OR_T[(unsigned char)in];
}
static void cyc_update(or_cyclic* c, char out, char in) {                /* :85-92 */
    uint64_t z = or_rotl(OR_T[(unsigned char)out], c->myr);
    c->h = or_rotl(c->h, 1) ^ z ^ OR_T[(unsigned char)in];
}
static void cyc_reverse_update(or_cyclic* c, char out, char in) {        /* :96-101 */
    uint64_t z = or_rotl(OR_T[(unsigned char)out], c->myr);
    c->h ^= z ^ OR_T[(unsigned char)in];
    c->h = or_rotr(c->h, 1);
}

/* LemireShifterPolicy<Canonical<uint64_t>>: rollinghashshifter.hh:38-49,128-219.
 * The forward shifter (FwdLemireShifter) is the fw half of the same state. */
typedef struct { or_cyclic fw, rc; int K; } or_shifter;

static void sh_init(or_shifter* s, int K) {
    s->K = K;
    s->fw.myr = (unsigned)(K % 64);   /* cyclichash.h:32 myr = n % wordsize */
    s->rc.myr = (unsigned)(K % 64);
    cyc_reset(&s->fw);
    cyc_reset(&s->rc);
}

/* rollinghashshifter.hh:180-197 hash_base_impl(const char*) */
static void sh_hash_base(or_shifter* s, const char* seq) {
    cyc_reset(&s->fw);
    cyc_reset(&s->rc);
    for (int i = 0; i < s->K; ++i) {
        cyc_eat(&s->fw, seq[i]);
        cyc_eat(&s->rc, or_complement(seq[s->K - i - 1]));
    }
}

/* rollinghashshifter.hh:200-208 shift_right_impl(out, in) */
static void sh_shift_right(or_shifter* s, char out, char in) {
    cyc_update(&s->fw, out, in);
    cyc_reverse_update(&s->rc, or_complement(in), or_complement(out));
}

/* rollinghashshifter.hh:211-219 shift_left_impl(in, out) */
static void sh_shift_left(or_shifter* s, char in, char out) {
    cyc_reverse_update(&s->fw, in, out);
    cyc_update(&s->rc, or_complement(out), or_complement(in));
}

/* canonical.hh:124-126 Canonical::value() = min(fw, rc); Hash::value() = fw. */
static uint64_t sh_value(const or_shifter* s, int canonical) {
    if (!canonical) return s->fw.h;
    return s->fw.h < s->rc.h ? s->fw.h : s->rc.h;
}

void or_hash_kmer(const char* s, int K, uint64_t* fw, uint64_t* rc) {
    or_shifter sh;
    sh_init(&sh, K);
    sh_hash_base(&sh, s);
    *fw = sh.fw.h;
    *rc = sh.rc.h;
}

/* KmerIterator<Shifter>: kmeriterator.hh:64-76 (hash_base on the first K
 * bytes) and :109-123 (next() = shift_right(seq[i-1], seq[i+K-1])). */
size_t or_hash_read(const char* s, size_t len, int K, uint64_t* fw, uint64_t* rc) {
    if (len < (size_t)K) return 0;
    or_shifter sh;
    sh_init(&sh, K);
    sh_hash_base(&sh, s);
    size_t n = len - (size_t)K + 1;
    fw[0] = sh.fw.h;
    rc[0] = sh.rc.h;
    for (size_t i = 1; i < n; ++i) {
        sh_shift_right(&sh, s[i - 1], s[i + K - 1]);
        fw[i] = sh.fw.h;
        rc[i] = sh.rc.h;
    }
    return n;
}

/* storage.hh:146-164 is_prime (trial division by odd i < sqrt(n)+1). */
static int or_is_prime(uint64_t n) {
    if (n < 2) return 0;
    if (n == 2) return 1;
    if (n % 2 == 0) return 0;
    for (unsigned long long i = 3; i < sqrt((double)n) + 1; i += 2) {
        if (n % i == 0) return 0;
    }
    return 1;
}

/* storage.hh:166-190 get_n_primes_near_x: odd candidates strictly below x,
 * descending, until n primes are found or the candidate reaches 1. */
uint32_t or_primes_near_x(uint32_t n, uint64_t x, uint64_t* out) {
    uint32_t found = 0;
    if (x == 1) {
        out[0] = 1;
        return 1;
    }
    uint64_t i = x - 1;
    if (i % 2 == 0) i--;
    while (found != n) {
        if (or_is_prime(i)) out[found++] = i;
        if (i == 1) break;
        i -= 2;
    }
    return found;
}

/* ---------------------------------------------------------------- storages */

/* bitstorage.hh:143-156 (p/8+1 bytes) and bytestorage.hh:125-134 (p bytes). */
uint64_t or_table_bytes(const or_storage* st, uint32_t i) {
    if (st->kind == OR_BIT) return st->sizes[i] / 8 + 1;
    return st->sizes[i];
}

static uint64_t or_mix(uint64_t k) {
    k ^= k >> 33; k *= 0xff51afd7ed558ccdULL;
    k ^= k >> 33; k *= 0xc4ceb9fe1a85ec53ULL;
    k ^= k >> 33;
    return k;
}

static const uint64_t OR_EMPTY = 0xFFFFFFFFFFFFFFFFULL;

int or_storage_init(or_storage* st, int kind, const uint64_t* sizes, uint32_t n_tables) {
    memset(st, 0, sizeof(*st));
    st->kind = kind;
    if (kind == OR_EXACT) {
        st->n_occupied = 32;   /* spp.h:2558 HT_DEFAULT_STARTING_BUCKETS */
        st->capacity = 1024;
        st->keys = (uint64_t*)malloc(st->capacity * sizeof(uint64_t));
        if (!st->keys) return -1;
        memset(st->keys, 0xFF, st->capacity * sizeof(uint64_t));
        return 0;
    }
    if (n_tables == 0 || n_tables > OR_MAX_TABLES) return -1;
    st->n_tables = n_tables;
    for (uint32_t i = 0; i < n_tables; ++i) {
        st->sizes[i] = sizes[i];
        st->tables[i] = (uint8_t*)calloc(or_table_bytes(st, i), 1);
        if (!st->tables[i]) return -1;
    }
    return 0;
}

void or_storage_free(or_storage* st) {
    for (uint32_t i = 0; i < st->n_tables; ++i) free(st->tables[i]);
    free(st->keys);
    free(st->big_keys);
    free(st->big_vals);
    memset(st, 0, sizeof(*st));
}

static void exact_grow(or_storage* st) {
    uint64_t old_cap = st->capacity;
    uint64_t* old = st->keys;
    st->capacity = old_cap * 2;
    st->keys = (uint64_t*)malloc(st->capacity * sizeof(uint64_t));
    memset(st->keys, 0xFF, st->capacity * sizeof(uint64_t));
    for (uint64_t j = 0; j < old_cap; ++j) {
        if (old[j] == OR_EMPTY) continue;
        uint64_t s = or_mix(old[j]) & (st->capacity - 1);
        while (st->keys[s] != OR_EMPTY) s = (s + 1) & (st->capacity - 1);
        st->keys[s] = old[j];
    }
    free(old);
}

/* sparseppstorage.hh:92-97 insert: true iff the value was absent. */
static int exact_insert(or_storage* st, uint64_t h) {
    if (h == OR_EMPTY) {
        int was = st->has_ones_key;
        st->has_ones_key = 1;
        if (!was) st->count++;
        return !was;
    }
    if (2 * (st->count + 1) > st->capacity) exact_grow(st);
    uint64_t s = or_mix(h) & (st->capacity - 1);
    while (st->keys[s] != OR_EMPTY) {
        if (st->keys[s] == h) return 0;
        s = (s + 1) & (st->capacity - 1);
    }
    st->keys[s] = h;
    st->count++;
    return 1;
}

static int exact_contains(const or_storage* st, uint64_t h) {
    if (h == OR_EMPTY) return st->has_ones_key;
    uint64_t s = or_mix(h) & (st->capacity - 1);
    while (st->keys[s] != OR_EMPTY) {
        if (st->keys[s] == h) return 1;
        s = (s + 1) & (st->capacity - 1);
    }
    return 0;
}

uint64_t or_exact_keys(const or_storage* st, uint64_t* out, uint64_t cap) {
    uint64_t n = 0;
    for (uint64_t j = 0; j < st->capacity; ++j) {
        if (st->keys[j] != OR_EMPTY) {
            if (n < cap) out[n] = st->keys[j];
            n++;
        }
    }
    if (st->has_ones_key) {
        if (n < cap) out[n] = OR_EMPTY;
        n++;
    }
    return n;
}

/* bitstorage.hh:195-219 BitStorage::insert */
static int bit_insert(or_storage* st, uint64_t h) {
    int is_new = 0;
    for (uint32_t i = 0; i < st->n_tables; ++i) {
        uint64_t bin = h % st->sizes[i];
        uint64_t byte = bin / 8;
        uint8_t bit = (uint8_t)(1u << (bin % 8));
        uint8_t orig = st->tables[i][byte];
        st->tables[i][byte] = (uint8_t)(orig | bit);
        if (!(orig & bit)) {
            if (i == 0) st->n_occupied++;
            is_new = 1;
        }
    }
    if (is_new) st->n_unique++;
    return is_new;
}

/* bitstorage.cc:87-100 BitStorage::query */
static int16_t bit_query(const or_storage* st, uint64_t h) {
    for (uint32_t i = 0; i < st->n_tables; ++i) {
        uint64_t bin = h % st->sizes[i];
        if (!(st->tables[i][bin / 8] & (1u << (bin % 8)))) return 0;
    }
    return 1;
}

/* bytestorage.cc:60-113 ByteStorage::insert (bigcount disabled, its default) */
/* The bigcount map: open addressing over keys (OR_EMPTY marks a free slot;
 * the all-ones key has its own flag), int16 values. */
static int16_t* big_slot(or_storage* st, uint64_t h, int create) {
    if (h == OR_EMPTY) {
        if (!st->big_has_ones) {
            if (!create) return NULL;
            st->big_has_ones = 1;
            st->big_ones_val = 0;
        }
        return &st->big_ones_val;
    }
    if (!st->big_keys) {
        if (!create) return NULL;
        st->big_cap = 64;
        st->big_keys = (uint64_t*)malloc(st->big_cap * sizeof(uint64_t));
        st->big_vals = (int16_t*)calloc(st->big_cap, sizeof(int16_t));
        memset(st->big_keys, 0xFF, st->big_cap * sizeof(uint64_t));
    }
    uint64_t s = or_mix(h) & (st->big_cap - 1);
    while (st->big_keys[s] != OR_EMPTY) {
        if (st->big_keys[s] == h) return &st->big_vals[s];
        s = (s + 1) & (st->big_cap - 1);
    }
    if (!create) return NULL;
    if (2 * (st->big_count + 1) > st->big_cap) {
        uint64_t old_cap = st->big_cap;
        uint64_t* ok = st->big_keys;
        int16_t* ov = st->big_vals;
        st->big_cap = old_cap * 2;
        st->big_keys = (uint64_t*)malloc(st->big_cap * sizeof(uint64_t));
        st->big_vals = (int16_t*)calloc(st->big_cap, sizeof(int16_t));
        memset(st->big_keys, 0xFF, st->big_cap * sizeof(uint64_t));
        for (uint64_t j = 0; j < old_cap; ++j) {
            if (ok[j] == OR_EMPTY) continue;
            uint64_t t = or_mix(ok[j]) & (st->big_cap - 1);
            while (st->big_keys[t] != OR_EMPTY) t = (t + 1) & (st->big_cap - 1);
            st->big_keys[t] = ok[j];
            st->big_vals[t] = ov[j];
        }
        free(ok);
        free(ov);
        return big_slot(st, h, 1);
    }
    st->big_keys[s] = h;
    st->big_vals[s] = 0;
    st->big_count++;
    return &st->big_vals[s];
}

void or_set_use_bigcount(or_storage* st, int b) { st->use_bigcount = b ? 1 : 0; }

int or_bigcount_get(const or_storage* st, uint64_t h, int16_t* v) {
    int16_t* p = big_slot((or_storage*)st, h, 0);
    if (!p) return 0;
    *v = *p;
    return 1;
}

void or_bigcount_set(or_storage* st, uint64_t h, int16_t v) { *big_slot(st, h, 1) = v; }

/* The map's pairs, sorted by key (the order of goetia's unordered_map is its
 * own); returns the number of pairs. */
static int cmp_u64(const void* a, const void* b) {
    uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? -1 : x > y;
}

uint64_t or_bigcount_pairs(const or_storage* st, uint64_t* keys, int16_t* vals, uint64_t cap) {
    uint64_t n = st->big_count + (st->big_has_ones ? 1 : 0);
    if (!keys || cap < n) return n;
    uint64_t k = 0;
    for (uint64_t j = 0; j < st->big_cap; ++j)
        if (st->big_keys[j] != OR_EMPTY) keys[k++] = st->big_keys[j];
    if (st->big_has_ones) keys[k++] = OR_EMPTY;
    qsort(keys, k, sizeof(uint64_t), cmp_u64);
    for (uint64_t j = 0; j < k; ++j) or_bigcount_get(st, keys[j], &vals[j]);
    return n;
}

/* bytestorage.cc:60-114 ByteStorage::insert.  A table whose count is already
 * 255 is "full"; when every table is full and bigcount is on, the k-mer's
 * map entry (0 when absent, as operator[] creates it) becomes 256 if it was
 * 0, else goes up by one while the int16 value compares below 65535 as an
 * unsigned int: 1..32766 step up, 32767 wraps to -32768, and a negative
 * value (a huge unsigned) stays. */
static int byte_insert(or_storage* st, uint64_t h) {
    int is_new = 0;
    uint32_t n_full = 0;
    for (uint32_t i = 0; i < st->n_tables; ++i) {
        uint64_t bin = h % st->sizes[i];
        uint8_t cur = st->tables[i][bin];
        if (!is_new) {
            if (cur == 0) {
                is_new = 1;
                if (i == 0) st->n_occupied++;
            }
        }
        if (cur < 255) st->tables[i][bin] = (uint8_t)(cur + 1);
        else n_full++;
    }
    if (n_full == st->n_tables && st->use_bigcount) {
        int16_t* v = big_slot(st, h, 1);
        if (*v == 0) *v = 256;
        else if ((unsigned)(int)*v < 65535u) *v = (int16_t)(uint16_t)((uint16_t)*v + 1u);
    }
    if (is_new) st->n_unique++;
    return is_new;
}

/* bytestorage.cc:116-139 ByteStorage::query: min over tables, bounded by 255 */
static int16_t byte_query(const or_storage* st, uint64_t h) {
    int16_t m = 255;
    for (uint32_t i = 0; i < st->n_tables; ++i) {
        int16_t c = st->tables[i][h % st->sizes[i]];
        if (c < m) m = c;
    }
    int16_t big;
    if (m == 255 && st->use_bigcount && or_bigcount_get(st, h, &big)) m = big;
    return m;
}

/* The exact set's n_occupied is sparsepp's bucket_count()
 * (sparseppstorage.hh:80-86).  Every insert first runs _resize_delta(1)
 * (spp.h:3223-3227): with no erasures, it keeps the buckets while
 * size + 1 <= enlarge_threshold = buckets * 0.5 (spp.h:2706-2710, 419-421,
 * HT_OCCUPANCY_PCT = 50 at :3542) and otherwise grows to
 * min_buckets(size + 1, buckets) (spp.h:2717-2723, 443-458). */
static uint64_t spp_min_buckets(uint64_t num_elts, uint64_t min_buckets_wanted) {
    uint64_t sz = 4;   /* HT_MIN_BUCKETS */
    while (sz < min_buckets_wanted || num_elts >= sz / 2) sz *= 2;
    return sz;
}

static void spp_resize_delta(or_storage* st) {
    uint64_t need = st->count + 1;
    if (st->n_occupied >= 4 && need <= st->n_occupied / 2) return;
    if (spp_min_buckets(need, 0) <= st->n_occupied) return;
    st->n_occupied = spp_min_buckets(need, st->n_occupied);
}

int or_insert(or_storage* st, uint64_t h) {
    switch (st->kind) {
        case OR_BIT:  return bit_insert(st, h);
        case OR_BYTE: return byte_insert(st, h);
        default: {
            spp_resize_delta(st);
            int n = exact_insert(st, h);
            if (n) st->n_unique++;
            return n;
        }
    }
}

int16_t or_query(const or_storage* st, uint64_t h) {
    switch (st->kind) {
        case OR_BIT:  return bit_query(st, h);
        case OR_BYTE: return byte_query(st, h);
        default:      return (int16_t)exact_contains(st, h);   /* sparseppstorage.cc:28-31 */
    }
}

/* bitstorage.cc:78-84 (always 1), bytestorage.cc:142-150 (1 if new, else
 * the post-insert query), sparseppstorage.cc:21-25 (always 1). */
int16_t or_insert_and_query(or_storage* st, uint64_t h) {
    int n = or_insert(st, h);
    if (st->kind == OR_BYTE) return n ? 1 : byte_query(st, h);
    return 1;
}

/* ------------------------------------------------------- sequence level */

static int is_acgt(char c) { return c == 'A' || c == 'C' || c == 'G' || c == 'T'; }

/* Fills valid[j] for the len-K+1 windows of s under the input policy. */
static size_t window_mask(const char* s, size_t len, int K, int mode, uint8_t* valid) {
    if (len < (size_t)K) return 0;
    size_t n = len - (size_t)K + 1;
    if (mode == OR_MODE_RAW) {
        memset(valid, 1, n);
        return n;
    }
    /* skip-N: a window is valid iff its K bytes are all A/C/G/T */
    size_t bad = 0;
    for (int j = 0; j < K; ++j) bad += !is_acgt(s[j]);
    valid[0] = bad == 0;
    for (size_t i = 1; i < n; ++i) {
        bad -= !is_acgt(s[i - 1]);
        bad += !is_acgt(s[i + K - 1]);
        valid[i] = bad == 0;
    }
    return n;
}

typedef struct { uint64_t* fw; uint64_t* rc; uint8_t* valid; size_t n; } or_windows;

static int windows_of(const char* s, size_t len, int K, int mode, or_windows* w) {
    w->n = 0;
    w->fw = w->rc = NULL;
    w->valid = NULL;
    if (len < (size_t)K) return 0;
    size_t n = len - (size_t)K + 1;
    w->fw = (uint64_t*)malloc(n * sizeof(uint64_t));
    w->rc = (uint64_t*)malloc(n * sizeof(uint64_t));
    w->valid = (uint8_t*)malloc(n);
    or_hash_read(s, len, K, w->fw, w->rc);
    window_mask(s, len, K, mode, w->valid);
    w->n = n;
    return 1;
}

static void windows_free(or_windows* w) { free(w->fw); free(w->rc); free(w->valid); }

static uint64_t wval(const or_windows* w, size_t i, int canonical) {
    if (!canonical) return w->fw[i];
    return w->fw[i] < w->rc[i] ? w->fw[i] : w->rc[i];
}

/* dbg.hh:307-318 insert_sequence(seq, n_new); under skip-N each ACGT run of
 * length >= K is inserted in position order, as repeated calls would. */
uint64_t or_insert_sequence(or_storage* st, int K, int canonical, int mode,
                            const char* s, size_t len, uint64_t* n_new, uint8_t* new_flags) {
    or_windows w;
    uint64_t hashed = 0, nn = 0;
    if (windows_of(s, len, K, mode, &w)) {
        for (size_t i = 0; i < w.n; ++i) {
            int f = 0;
            if (w.valid[i]) {
                f = or_insert(st, wval(&w, i, canonical));
                nn += (uint64_t)f;
                hashed++;
            }
            if (new_flags) new_flags[i] = (uint8_t)f;
        }
        windows_free(&w);
    }
    if (n_new) *n_new = nn;
    return hashed;
}

/* dbg.hh:327-340 insert_and_query_sequence */
uint64_t or_insert_and_query_sequence(or_storage* st, int K, int canonical, int mode,
                                      const char* s, size_t len, int16_t* counts) {
    or_windows w;
    uint64_t hashed = 0;
    if (windows_of(s, len, K, mode, &w)) {
        for (size_t i = 0; i < w.n; ++i) {
            if (w.valid[i]) {
                counts[i] = or_insert_and_query(st, wval(&w, i, canonical));
                hashed++;
            } else {
                counts[i] = -1;
            }
        }
        windows_free(&w);
    }
    return hashed;
}

/* dbg.hh:364-376 query_sequence(seq, counts, hashes) */
uint64_t or_query_sequence(const or_storage* st, int K, int canonical, int mode,
                           const char* s, size_t len, int16_t* counts, uint64_t* hashes) {
    or_windows w;
    uint64_t hashed = 0;
    if (windows_of(s, len, K, mode, &w)) {
        for (size_t i = 0; i < w.n; ++i) {
            if (w.valid[i]) {
                uint64_t h = wval(&w, i, canonical);
                if (counts) counts[i] = or_query(st, h);
                if (hashes) hashes[i] = h;
                hashed++;
            } else {
                if (counts) counts[i] = -1;
                if (hashes) hashes[i] = 0;
            }
        }
        windows_free(&w);
    }
    return hashed;
}

/* Neighbour extensions of one k-mer, computed with the shift operations the
 * reference uses: hashextender.hh:352-364 left_extensions_impl (shift_left by
 * each of "ACGT", alphabets.hh:109, then undo with shift_right) and :386-396
 * right_extensions_impl; counted as in unitig_walker.hh:322-330,361-371
 * (present iff query >= min_count; the reference's test is query != 0,
 * i.e. min_count 1). */
void or_kmer_degrees(const or_storage* st, int K, int canonical, int min_count,
                     const char* kmer, uint8_t* in_deg, uint8_t* out_deg,
                     uint64_t left[4], uint64_t right[4]) {
    static const char SYM[4] = {'A', 'C', 'G', 'T'};
    or_shifter sh;
    sh_init(&sh, K);
    sh_hash_base(&sh, kmer);
    char front = kmer[0], back = kmer[K - 1];
    uint8_t nin = 0, nout = 0;
    for (int c = 0; c < 4; ++c) {
        sh_shift_left(&sh, SYM[c], back);
        uint64_t v = sh_value(&sh, canonical);
        if (left) left[c] = v;
        if (or_query(st, v) >= min_count) nin++;
        sh_shift_right(&sh, SYM[c], back);
    }
    for (int c = 0; c < 4; ++c) {
        sh_shift_right(&sh, front, SYM[c]);
        uint64_t v = sh_value(&sh, canonical);
        if (right) right[c] = v;
        if (or_query(st, v) >= min_count) nout++;
        sh_shift_left(&sh, front, SYM[c]);
    }
    *in_deg = nin;
    *out_deg = nout;
}

/* Per-window in/out degrees of a read, as find_decision_kmers visits them
 * (unitig_walker.hh:1025-1045 with get_decision_neighbors :1101-1117).
 * Skipped windows get degree 0xFF. */
uint64_t or_degrees_sequence(const or_storage* st, int K, int canonical, int mode,
                             int min_count, const char* s, size_t len,
                             uint8_t* in_deg, uint8_t* out_deg) {
    if (len < (size_t)K) return 0;
    size_t n = len - (size_t)K + 1;
    uint8_t* valid = (uint8_t*)malloc(n);
    window_mask(s, len, K, mode, valid);
    uint64_t done = 0;
    for (size_t i = 0; i < n; ++i) {
        if (!valid[i]) {
            in_deg[i] = out_deg[i] = 0xFF;
            continue;
        }
        or_kmer_degrees(st, K, canonical, min_count, s + i, &in_deg[i], &out_deg[i], NULL, NULL);
        done++;
    }
    free(valid);
    return done;
}

/* FileProcessor::advance order (processors.hh:208-229): reads in order. */
uint64_t or_insert_reads(or_storage* st, int K, int canonical, int mode,
                         const char* seq, const int64_t* offsets, uint64_t n_reads,
                         uint64_t* n_new) {
    uint64_t total = 0;
    for (uint64_t r = 0; r < n_reads; ++r) {
        uint64_t nn = 0;
        total += or_insert_sequence(st, K, canonical, mode, seq + offsets[r],
                                    (size_t)(offsets[r + 1] - offsets[r]), &nn, NULL);
        if (n_new) n_new[r] = nn;
    }
    return total;
}

/* or_insert_reads with each window's first-touch flag, packed
 * max(0, len - K + 1) per read. */
uint64_t or_insert_reads_flags(or_storage* st, int K, int canonical, int mode,
                               const char* seq, const int64_t* offsets, uint64_t n_reads,
                               uint64_t* n_new, uint8_t* new_flags) {
    uint64_t total = 0, pos = 0;
    for (uint64_t r = 0; r < n_reads; ++r) {
        uint64_t nn = 0;
        const size_t len = (size_t)(offsets[r + 1] - offsets[r]);
        total += or_insert_sequence(st, K, canonical, mode, seq + offsets[r], len, &nn,
                                    new_flags ? new_flags + pos : NULL);
        if (n_new) n_new[r] = nn;
        if (len >= (size_t)K) pos += len - (size_t)K + 1;
    }
    return total;
}

/* insert_and_query_sequence over reads in order (dbg.hh:327-340), counts
 * packed per read. */
uint64_t or_insert_and_query_reads(or_storage* st, int K, int canonical, int mode,
                                   const char* seq, const int64_t* offsets, uint64_t n_reads,
                                   int16_t* counts) {
    uint64_t total = 0, pos = 0;
    for (uint64_t r = 0; r < n_reads; ++r) {
        const size_t len = (size_t)(offsets[r + 1] - offsets[r]);
        if (len < (size_t)K) continue;
        total += or_insert_and_query_sequence(st, K, canonical, mode, seq + offsets[r], len, counts + pos);
        pos += len - (size_t)K + 1;
    }
    return total;
}

/* The value hashed for every window (min(fw, rc) or fw; 0 where skipped),
 * packed per read. */
uint64_t or_hash_reads(int K, int canonical, int mode, const char* seq, const int64_t* offsets,
                       uint64_t n_reads, uint64_t* out) {
    uint64_t pos = 0;
    for (uint64_t r = 0; r < n_reads; ++r) {
        or_windows w;
        const size_t len = (size_t)(offsets[r + 1] - offsets[r]);
        if (!windows_of(seq + offsets[r], len, K, mode, &w)) continue;
        for (size_t i = 0; i < w.n; ++i) out[pos + i] = w.valid[i] ? wval(&w, i, canonical) : 0;
        pos += w.n;
        windows_free(&w);
    }
    return pos;
}

/* DiginormFilter<dBG<ByteStorage, Sh>>::Filter::filter_sequence
 * (diginorm.hh:105-119) for one read: -1 when the read is shorter than K (the
 * k-mer iterator raises SequenceLengthException, kmeriterator.hh:72-74), 0
 * when median_count_at_least (diginorm.hh:39-76) holds, else 1 after the read
 * is inserted (dbg.hh:249-265).  median_count_at_least queries the first
 * min_req k-mers, min_req = (unsigned)(0.5 + float(n_kmers) / 2)
 * (diginorm.hh:43), counting those with count >= cutoff; it holds if all of
 * them did, else it goes on through the remaining k-mers until the count
 * reaches min_req (diginorm.hh:47-73).  Under skip-N, windows holding a
 * non-ACGT byte are never counted. */
int or_diginorm_filter(or_storage* st, int K, int canonical, int mode, unsigned cutoff,
                       const char* s, size_t len) {
    if (len < (size_t)K) return -1;
    or_windows w;
    windows_of(s, len, K, mode, &w);
    const unsigned min_req = (unsigned)(0.5 + (double)((float)(len - (size_t)K + 1) / 2));
    unsigned num = 0;
    int at_least = 0;
    size_t i = 0;
    for (; i < min_req; ++i)
        if (w.valid[i] && (unsigned)or_query(st, wval(&w, i, canonical)) >= cutoff) ++num;
    if (num >= min_req) {
        at_least = 1;
    } else {
        for (; i < w.n; ++i) {
            if (w.valid[i] && (unsigned)or_query(st, wval(&w, i, canonical)) >= cutoff) {
                ++num;
                if (num >= min_req) {
                    at_least = 1;
                    break;
                }
            }
        }
    }
    windows_free(&w);
    if (at_least) return 0;
    or_insert_sequence(st, K, canonical, mode, s, len, NULL, NULL);
    return 1;
}

/* The filter over reads in order: passed[r] = 1 for inserted reads, 0 for
 * dropped and short ones.  Returns the reads that passed. */
uint64_t or_diginorm_reads(or_storage* st, int K, int canonical, int mode, unsigned cutoff,
                           const char* seq, const int64_t* offsets, uint64_t n_reads, uint8_t* passed) {
    uint64_t n = 0;
    for (uint64_t r = 0; r < n_reads; ++r) {
        int p = or_diginorm_filter(st, K, canonical, mode, cutoff, seq + offsets[r],
                                   (size_t)(offsets[r + 1] - offsets[r]));
        passed[r] = (uint8_t)(p == 1);
        n += p == 1;
    }
    return n;
}

/* ------------------------------------------------- all host cores */

/* One window into Bloom / count-min tables with atomic table updates, as the
 * reference's storages do with __sync builtins (bitstorage.hh:195-219,
 * bytestorage.cc:60-113): OR of a bit, saturating increment of a byte by
 * compare-and-swap.  Both commute, so after any interleaving the tables equal
 * a sequential insert of the same windows; the counters (which depend on the
 * order) are not kept here. */
static void atomic_insert(or_storage* st, uint64_t h) {
    for (uint32_t i = 0; i < st->n_tables; ++i) {
        uint64_t bin = h % st->sizes[i];
        if (st->kind == OR_BIT) {
            __atomic_fetch_or(&st->tables[i][bin / 8], (uint8_t)(1u << (bin % 8)), __ATOMIC_RELAXED);
        } else {
            uint8_t* c = &st->tables[i][bin];
            uint8_t cur = __atomic_load_n(c, __ATOMIC_RELAXED);
            while (cur < 255 &&
                   !__atomic_compare_exchange_n(c, &cur, (uint8_t)(cur + 1), 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
            }
        }
    }
}

typedef struct {
    or_storage* st;
    int K, canonical, mode;
    const char* seq;
    const int64_t* offsets;
    const int64_t* kmer_off;     /* query only */
    int16_t* counts;             /* query only */
    uint64_t r0, r1;
    uint64_t hashed;
} or_mt_part;

static void* mt_insert(void* arg) {
    or_mt_part* a = (or_mt_part*)arg;
    for (uint64_t r = a->r0; r < a->r1; ++r) {
        or_windows w;
        if (!windows_of(a->seq + a->offsets[r], (size_t)(a->offsets[r + 1] - a->offsets[r]), a->K, a->mode, &w))
            continue;
        for (size_t i = 0; i < w.n; ++i)
            if (w.valid[i]) {
                atomic_insert(a->st, wval(&w, i, a->canonical));
                a->hashed++;
            }
        windows_free(&w);
    }
    return NULL;
}

static void* mt_query(void* arg) {
    or_mt_part* a = (or_mt_part*)arg;
    for (uint64_t r = a->r0; r < a->r1; ++r)
        a->hashed += or_query_sequence(a->st, a->K, a->canonical, a->mode, a->seq + a->offsets[r],
                                       (size_t)(a->offsets[r + 1] - a->offsets[r]), a->counts + a->kmer_off[r], NULL);
    return NULL;
}

typedef struct {
    or_mt_part base;
    int min_count;
    const uint8_t* only;
    uint8_t* in_deg;
    uint8_t* out_deg;
} or_mt_deg;

/* or_degrees_sequence for the windows with only[w] set (all when only is
 * NULL); 0xFF elsewhere. */
static void* mt_degrees(void* arg) {
    or_mt_deg* d = (or_mt_deg*)arg;
    or_mt_part* a = &d->base;
    for (uint64_t r = a->r0; r < a->r1; ++r) {
        const char* s = a->seq + a->offsets[r];
        const size_t len = (size_t)(a->offsets[r + 1] - a->offsets[r]);
        if (len < (size_t)a->K) continue;
        const size_t n = len - (size_t)a->K + 1;
        const int64_t k0 = a->kmer_off[r];
        uint8_t* valid = (uint8_t*)malloc(n);
        window_mask(s, len, a->K, a->mode, valid);
        for (size_t i = 0; i < n; ++i) {
            if (!valid[i] || (d->only && !d->only[k0 + i])) {
                d->in_deg[k0 + i] = d->out_deg[k0 + i] = 0xFF;
                continue;
            }
            or_kmer_degrees(a->st, a->K, a->canonical, d->min_count, s + i, &d->in_deg[k0 + i],
                            &d->out_deg[k0 + i], NULL, NULL);
            a->hashed++;
        }
        free(valid);
    }
    return NULL;
}

/* Reads split into n_threads ranges of about equal bases, each on its own
 * thread. */
static uint64_t run_parts(void* (*fn)(void*), or_mt_part* proto, uint64_t n_reads, int n_threads) {
    if (n_threads < 1) n_threads = 1;
    or_mt_part* parts = (or_mt_part*)calloc((size_t)n_threads, sizeof(or_mt_part));
    pthread_t* th = (pthread_t*)calloc((size_t)n_threads, sizeof(pthread_t));
    const int64_t total = n_reads ? proto->offsets[n_reads] : 0;
    uint64_t r = 0;
    for (int t = 0; t < n_threads; ++t) {
        parts[t] = *proto;
        parts[t].r0 = r;
        const int64_t target = total * (t + 1) / n_threads;
        while (r < n_reads && (t == n_threads - 1 || proto->offsets[r] < target)) ++r;
        parts[t].r1 = r;
        parts[t].hashed = 0;
    }
    for (int t = 1; t < n_threads; ++t) pthread_create(&th[t], NULL, fn, &parts[t]);
    fn(&parts[0]);
    uint64_t hashed = parts[0].hashed;
    for (int t = 1; t < n_threads; ++t) {
        pthread_join(th[t], NULL);
        hashed += parts[t].hashed;
    }
    free(parts);
    free(th);
    return hashed;
}

/* insert_sequence over reads on n_threads threads (Bloom / count-min only;
 * the tables end as a sequential insert leaves them, the counters are not
 * updated).  Returns the k-mers hashed. */
uint64_t or_insert_reads_mt(or_storage* st, int K, int canonical, int mode, const char* seq,
                            const int64_t* offsets, uint64_t n_reads, int n_threads) {
    if (st->kind != OR_BIT && st->kind != OR_BYTE) return 0;
    or_mt_part proto = {st, K, canonical, mode, seq, offsets, NULL, NULL, 0, 0, 0};
    return run_parts(mt_insert, &proto, n_reads, n_threads);
}

/* query_sequence over reads on n_threads threads, counts packed per read as
 * or_query_reads packs them. */
uint64_t or_query_reads_mt(const or_storage* st, int K, int canonical, int mode, const char* seq,
                           const int64_t* offsets, uint64_t n_reads, int16_t* counts, int n_threads) {
    int64_t* ko = (int64_t*)malloc((n_reads + 1) * sizeof(int64_t));
    ko[0] = 0;
    for (uint64_t r = 0; r < n_reads; ++r) {
        const int64_t len = offsets[r + 1] - offsets[r];
        ko[r + 1] = ko[r] + (len >= K ? len - K + 1 : 0);
    }
    or_mt_part proto = {(or_storage*)st, K, canonical, mode, seq, offsets, ko, counts, 0, 0, 0};
    const uint64_t hashed = run_parts(mt_query, &proto, n_reads, n_threads);
    free(ko);
    return hashed;
}

/* Counts are packed per read: read r owns max(0, len_r - K + 1) entries. */
uint64_t or_query_reads(const or_storage* st, int K, int canonical, int mode,
                        const char* seq, const int64_t* offsets, uint64_t n_reads,
                        int16_t* counts) {
    uint64_t total = 0, pos = 0;
    for (uint64_t r = 0; r < n_reads; ++r) {
        size_t len = (size_t)(offsets[r + 1] - offsets[r]);
        total += or_query_sequence(st, K, canonical, mode, seq + offsets[r], len,
                                   counts ? counts + pos : NULL, NULL);
        if (len >= (size_t)K) pos += len - (size_t)K + 1;
    }
    return total;
}

/* or_degrees_sequence over reads on n_threads threads, for the windows with
 * only[w] set (every window when only is NULL); in/out degrees packed per
 * read, 0xFF where skipped or not selected.  The storage is only read. */
uint64_t or_degrees_reads_mt(const or_storage* st, int K, int canonical, int mode, int min_count,
                             const char* seq, const int64_t* offsets, uint64_t n_reads, const uint8_t* only,
                             uint8_t* in_deg, uint8_t* out_deg, int n_threads) {
    if (n_threads < 1) n_threads = 1;
    int64_t* ko = (int64_t*)malloc((n_reads + 1) * sizeof(int64_t));
    ko[0] = 0;
    for (uint64_t r = 0; r < n_reads; ++r) {
        const int64_t len = offsets[r + 1] - offsets[r];
        ko[r + 1] = ko[r] + (len >= K ? len - K + 1 : 0);
    }
    or_mt_deg* parts = (or_mt_deg*)calloc((size_t)n_threads, sizeof(or_mt_deg));
    pthread_t* th = (pthread_t*)calloc((size_t)n_threads, sizeof(pthread_t));
    const int64_t total = n_reads ? offsets[n_reads] : 0;
    uint64_t r = 0;
    for (int t = 0; t < n_threads; ++t) {
        or_mt_part p = {(or_storage*)st, K, canonical, mode, seq, offsets, ko, NULL, r, 0, 0};
        const int64_t target = total * (t + 1) / n_threads;
        while (r < n_reads && (t == n_threads - 1 || offsets[r] < target)) ++r;
        p.r1 = r;
        parts[t].base = p;
        parts[t].min_count = min_count;
        parts[t].only = only;
        parts[t].in_deg = in_deg;
        parts[t].out_deg = out_deg;
    }
    for (int t = 1; t < n_threads; ++t) pthread_create(&th[t], NULL, mt_degrees, &parts[t]);
    mt_degrees(&parts[0]);
    uint64_t done = parts[0].base.hashed;
    for (int t = 1; t < n_threads; ++t) {
        pthread_join(th[t], NULL);
        done += parts[t].base.hashed;
    }
    free(parts);
    free(th);
    free(ko);
    return done;
}
