// C++ facade check: goetia's dBG template API over libgoetia_amd (GPU).
// Built by tests/test_cpp_facade.py with the host compiler against
// include/ and goetia_amd/libgoetia_amd.so; prints "FACADE OK" on success.
#include <cassert>
#include <cstdio>
#include <fstream>
#include <set>
#include <string>
#include <vector>

#include "goetia/dbg.hh"
#include "goetia/parsing/readers.hh"
#include "goetia/processors.hh"
#include "goetia/solidifier.hh"
#include "goetia/diginorm.hh"

using namespace goetia;

#define CHECK(c)                                                     \
    do {                                                             \
        if (!(c)) {                                                  \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            return 1;                                                \
        }                                                            \
    } while (0)

template <class S>
int run_presence(std::shared_ptr<S> store) {
    auto g = dBG<S, CanLemireShifter>::build(store, 21);
    const std::string s = "ACGTTGCAAGGCTTACCGATAGCATTACGGATCCAGT";
    for (size_t i = 0; i + 21 <= s.size(); ++i) {
        const std::string k = s.substr(i, 21);
        CHECK(g->query(k) == 0);
        CHECK(g->insert(k) == true);
        CHECK(g->query(k) == 1);
        CHECK(g->insert(k) == false);
    }
    return 0;
}

int main(int argc, char** argv) {
    // known-answer vector (goetia tests/test_hashing.py:15-18)
    CanLemireShifter sh(27);
    auto h = sh.hash_base("TCACCTGTGTTGTGCTACTTGCGGCGC");
    CHECK(h.fw_hash == 13194817695400542713ULL);
    CHECK(h.rc_hash == 4324216031038051805ULL);
    CHECK(FwdLemireShifter::hash("TCACCTGTGTTGTGCTACTTGCGGCGC", 27).value() == 13194817695400542713ULL);

    // KmerIterator (GPU) == per-k-mer host hashes
    const std::string seq = "TCACCTGTGTTGTGCTACTTGCGGCGCAAGTCCATGATCGATCGGGTAC";
    KmerIterator<CanLemireShifter> it(seq, 21);
    CanLemireShifter sh21(21);
    for (size_t i = 0; !it.done(); ++i) CHECK(it.next() == sh21.hash(seq.substr(i, 21)));

    CHECK(run_presence(BitStorage::build(100000, 4)) == 0);
    CHECK(run_presence(SparseppSetStorage::build()) == 0);

    // counting storage: counts after repeated inserts, insert_and_query
    auto cs = ByteStorage::build(100000, 4);
    CHECK(cs->get_tablesizes()[0] == 99991);
    auto g = dBG<ByteStorage, CanLemireShifter>::build(cs, 21);
    for (int it2 = 0; it2 < 5; ++it2) {
        auto c = g->query_sequence(seq);
        for (auto v : c) CHECK(v == it2);
        g->insert_sequence(seq);
    }
    std::vector<CanLemireShifter::hash_type> hashes;
    std::vector<count_t> counts;
    g->insert_sequence(seq, hashes, counts);
    CHECK(hashes.size() == seq.size() - 20);
    for (auto v : counts) CHECK(v == 6);

    // per-read n_new and batched calls
    auto g2 = dBG<BitStorage, CanLemireShifter>::build(BitStorage::build(100000, 4), 21);
    uint64_t n_new = 0;
    CHECK(g2->insert_sequence(seq, n_new) == seq.size() - 20);
    CHECK(n_new == g2->n_unique());
    ReadBatch batch({seq, seq.substr(5), "ACGT"});
    auto nn = g2->insert_sequences(batch);
    CHECK(nn.size() == 3 && nn[0] == 0 && nn[1] == 0 && nn[2] == 0);
    auto q = g2->query_sequences(batch);
    for (auto v : q) CHECK(v == 1);

    // neighbours on a linear path (test_dbg.py:182-196)
    auto g3 = dBG<SparseppSetStorage, CanLemireShifter>::build(SparseppSetStorage::build(), 21);
    const std::string path = seq.substr(0, 23);
    g3->insert_sequence(path);
    auto nb = g3->neighbors(path.substr(1, 21));
    CHECK(nb.first.size() == 1 && nb.second.size() == 1);
    CHECK(!g3->is_decision_kmer(path.substr(1, 21)));
    // a right fork: two right neighbour k-mers of the fork k-mer
    {
        auto g9 = dBG<SparseppSetStorage, CanLemireShifter>::build(SparseppSetStorage::build(), 21);
        const std::string core = seq.substr(0, 40);
        const char a = seq[40], b = a == 'A' ? 'C' : 'A';
        g9->insert_sequence(core + a + seq.substr(41, 20));
        g9->insert_sequence(core + b + seq.substr(41, 20));
        const std::string fork = core.substr(core.size() - 21);
        dBG<SparseppSetStorage, CanLemireShifter>::neighbor_pair_type nbp;
        CHECK(g9->get_decision_neighbors(fork, nbp));
        CHECK(nbp.first.size() == 1 && nbp.second.size() == 2);
        std::set<std::string> rk{nbp.second[0].kmer, nbp.second[1].kmer};
        CHECK(rk.count(fork.substr(1) + a) && rk.count(fork.substr(1) + b));
        std::vector<uint32_t> dpos;
        std::vector<dBG<SparseppSetStorage, CanLemireShifter>::hash_type> dh;
        std::vector<dBG<SparseppSetStorage, CanLemireShifter>::neighbor_pair_type> dn;
        g9->find_decision_kmers(core + a + seq.substr(41, 20), dpos, dh, dn);
        CHECK(dpos.size() == dn.size() && !dpos.empty());
        // cursor walk (hashextender.hh:327-353)
        g9->set_cursor(core.substr(0, 21));
        for (size_t i = 1; i + 21 <= core.size(); ++i) {
            CHECK(g9->shift_right(core[i + 20]).value() == g9->hash(core.substr(i, 21)).value());
            CHECK(g9->get_cursor() == core.substr(i, 21));
        }
        CHECK(g9->shift_left(core[core.size() - 22]).value() == g9->hash(core.substr(core.size() - 22, 21)).value());
    }

    // errors
    bool threw = false;
    try {
        g3->insert_sequence("ACGT");
    } catch (const SequenceLengthException&) {
        threw = true;
    }
    CHECK(threw);

    // clone() is empty; get_raw() matches the device tables' layout
    auto c2 = g->clone();
    CHECK(c2->n_unique() == 0);
    CHECK(g->get_raw() != nullptr);

    // FASTX input and the chunk driver (argv[1]: goetia's test-fastq-reads.fq,
    // argv[2]: a scratch output file)
    if (argc > 1) {
        auto parser = FastxParser<>::build(argv[1]);
        uint64_t n_rec = 0, kmers = 0, expected_stops = 0;
        IntervalCounter timer(1000);
        while (auto rec = parser->next()) {
            ++n_rec;
            const uint64_t k = rec->sequence.size() >= 21 ? rec->sequence.size() - 20 : 0;
            kmers += k;
            expected_stops += timer.poll(k);
        }
        CHECK(n_rec == parser->n_parsed() && n_rec > 0 && parser->n_skipped() == 0);
        auto g4 = dBG<ByteStorage, CanLemireShifter>::build(ByteStorage::build(100000, 4), 21);
        auto proc = InserterProcessor<dBG<ByteStorage, CanLemireShifter>>::build(g4, 1000);
        auto p2 = FastxParser<>::build(argv[1]);
        uint64_t stops = 0, n_seq = 0, t = 0;
        bool remaining = true;
        while (remaining) {
            std::tie(n_seq, t, remaining) = proc->advance(p2);
            stops += remaining;
        }
        CHECK(n_seq == n_rec && t == kmers && stops == expected_stops && expected_stops > 0);
        auto g5 = dBG<ByteStorage, CanLemireShifter>::build(ByteStorage::build(100000, 4), 21);
        auto p3 = FastxParser<>::build(argv[1]);
        for (auto b = p3->next_batch(); b.size(); b = p3->next_batch()) g5->insert_sequences(b);
        CHECK(g4->n_unique() == g5->n_unique() && g4->n_occupied() == g5->n_occupied());
        // solid filter: every read of a fresh graph is new, so with threshold 2 none passes,
        // and the second time through every read passes
        auto g6 = dBG<ByteStorage, CanLemireShifter>::build(ByteStorage::build(100000, 4), 21);
        auto sf = StreamingSolidFilter<dBG<ByteStorage, CanLemireShifter>>::Filter::build(g6, 0.75, 2);
        auto p4 = FastxParser<>::build(argv[1]);
        auto b4 = p4->next_batch();
        auto first = sf->filter_sequences(b4);
        auto second = sf->filter_sequences(b4);
        uint64_t n1 = 0, n2 = 0;
        for (auto v : first) n1 += v;
        for (auto v : second) n2 += v;
        CHECK(n2 == b4.size() && n1 < n2);
        // diginorm with cutoff 1: every read of the batch passes the first
        // time (its k-mers are new), and a second pass over the same reads
        // drops every read (all its k-mers now have count >= 1)
        auto g7 = dBG<ByteStorage, CanLemireShifter>::build(ByteStorage::build(1000000, 4), 21);
        auto dn = DiginormFilter<dBG<ByteStorage, CanLemireShifter>>::Filter::build(g7, 1);
        auto p5 = FastxParser<>::build(argv[1]);
        auto b5 = p5->next_batch();
        uint64_t d1 = 0, d2 = 0;
        for (auto v : dn->filter_sequences(b5)) d1 += v;
        for (auto v : dn->filter_sequences(b5)) d2 += v;
        CHECK(d1 > 0 && d2 == 0);
        using DF = DiginormFilter<dBG<ByteStorage, CanLemireShifter>>;
        CHECK(DF::median_count_at_least(b5.read(0), 1, *g7));
        CHECK(!std::get<0>(dn->filter_sequence(b5.read(0))));
        // FilterProcessor: the reads the filter passes, written as FASTQ
        auto g8 = dBG<ByteStorage, CanLemireShifter>::build(ByteStorage::build(1000000, 4), 21);
        auto dn8 = DF::Filter::build(g8, 1);
        const std::string out = argc > 2 ? std::string(argv[2]) : std::string(argv[1]) + ".facade_out.fq";
        uint64_t n_proc = 0, n_pass = 0;
        {
            auto fp = DF::Processor::build(dn8, out, 1000);
            auto p6 = FastxParser<>::build(argv[1]);
            uint64_t t8 = 0, stops8 = 0;
            bool rem = true;
            while (rem) {
                std::tie(n_proc, t8, rem) = fp->advance(p6);
                stops8 += rem;
            }
            n_pass = fp->n_passed();
            CHECK(n_proc == n_rec && t8 == kmers && stops8 == expected_stops);
        }
        uint64_t lines = 0, d3 = 0;
        {
            std::ifstream in(out);
            std::string line;
            while (std::getline(in, line)) ++lines;
        }
        for (auto v : dn8->filter_sequences(b5)) d3 += v;
        CHECK(n_pass == d1 && lines == 4 * n_pass && d3 == 0);
        std::remove(out.c_str());
    }

    // window_neighbors (GPU) == the cursor's left / right extensions and the
    // union-node filter (host shifts, per-hash queries)
    {
        auto gw = dBG<ByteStorage, CanLemireShifter>::build(ByteStorage::build(100000, 4), 21);
        gw->insert_sequence(seq);
        const std::string other = seq.substr(3, 30) + "ACGTTGCA";
        ReadBatch wb({seq, other});
        std::vector<uint8_t> in_deg, out_deg, present;
        std::vector<uint64_t> vals;
        gw->window_neighbors(wb, in_deg, out_deg, vals, present);
        size_t k = 0;
        for (const std::string& r : {seq, other}) {
            for (size_t i = 0; i + 21 <= r.size(); ++i, ++k) {
                gw->set_cursor(r.substr(i, 21));
                auto l = gw->left_extensions();
                auto rt = gw->right_extensions();
                uint8_t bits = 0;
                for (int c = 0; c < 4; ++c) {
                    CHECK(vals[8 * k + c] == l[c].value());
                    CHECK(vals[8 * k + 4 + c] == rt[c].value());
                }
                for (auto& e : gw->filter_nodes(l))
                    bits |= (uint8_t)(1u << (std::string("ACGT").find(e.symbol)));
                for (auto& e : gw->filter_nodes(rt))
                    bits |= (uint8_t)(1u << (4 + std::string("ACGT").find(e.symbol)));
                CHECK(present[k] == bits);
                CHECK(in_deg[k] == __builtin_popcount(bits & 15u) && out_deg[k] == __builtin_popcount(bits >> 4));
                std::set<CanLemireShifter::hash_type> extra{l[0].hash, rt[3].hash};
                CHECK(gw->in_degree(extra) == gw->filter_nodes(l, extra).size());
                CHECK(gw->out_degree(extra) == (size_t)__builtin_popcount((bits >> 4) | 8u));
            }
        }
        CHECK(k == vals.size() / 8);
    }
    std::printf("FACADE OK\n");
    return 0;
}
