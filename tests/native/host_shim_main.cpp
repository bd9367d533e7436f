// Drives the C++ host classes (hero_amd/host/hero_host.h) the way the reference's main.cpp drives
// its own: used by tests/test_gpu_host_shims.py.
//   host_shim_main correct <reads.fasta> <k> <m%> <i%> <l> <g|a> <out.fasta> [subject_chunk]
//   host_shim_main aln     <reads.fasta> <k> <m%> <i%> <l> <g|a> <out.aln>   [subject_chunk]
//   host_shim_main band    <g|a> <cases.txt> <out.txt>   (lines "S Q band flag" -> "band cigar")
// subject_chunk > 0 streams the subjects through SubjectDataset chunks instead of self-alignment.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../hero_amd/host/hero_host.h"

using namespace hero_host;

// correct: -c, corrected FASTA; aln: the .aln text; mapping: -c --mapping, the coverage histogram
// file (main.cpp:82-87 writes it instead of calling consensus)
static int run_path(bool correct, bool mapping, int argc, char **argv) {
	if (argc < 9) return 2;
	ReadBatch reads;
	if (!ReadBatch::loadFasta(argv[2], reads)) return 3;
	Config cfg;
	cfg.setKmer(std::atoi(argv[3]));
	cfg.setMaxMismatchRatePercent(std::atof(argv[4]));
	cfg.setMaxIndelRatePercent(std::atof(argv[5]));
	cfg.setMinimumOverlap(std::atoll(argv[6]));
	cfg.setAffine(argv[7][0] == 'a');
	cfg.setCallConsensus(correct);
	cfg.setMapping(mapping);
	const int64_t chunk = argc > 9 ? std::atoll(argv[9]) : 0;
	Device dev(cfg);
	QueryDataset q(dev, reads);
	HashTableLongKey table(&q);
	table.createHashTables();
	table.insertQueryDataset(&q);
	SubjectReadAligner aligner(&q, &table);
	if (chunk > 0) {
		SubjectDataset subjects(reads, chunk);
		aligner.start(subjects);
	} else {
		aligner.startSelf();
	}
	if (mapping) {
		std::ofstream(argv[8], std::ios_base::binary | std::ios_base::trunc).close();
		q.getCoverageAndWriteResults(argv[8]);
		return 0;
	}
	std::ofstream out(argv[8], std::ios_base::binary | std::ios_base::trunc);
	if (correct) {
		q.callConsensus();
		out << q.queryOutText();
	} else {
		out << aligner.alignments();
	}
	return 0;
}

static int run_band(int argc, char **argv) {
	if (argc < 5) return 2;
	Config cfg;
	cfg.setAffine(argv[2][0] == 'a');
	Device dev(cfg);
	BandAligner ba(dev);
	std::ifstream in(argv[3]);
	std::ofstream out(argv[4], std::ios_base::binary | std::ios_base::trunc);
	std::string line;
	while (std::getline(in, line)) {
		std::istringstream ls(line);
		std::string S, Q;
		int64_t band;
		int flag;
		if (!(ls >> S >> Q >> band >> flag)) continue;
		if (S == "-") S.clear();
		if (Q == "-") Q.clear();
		ba.setBand(band);
		ba.setSequences(S, 0, (int64_t) S.size(), Q, 0, (int64_t) Q.size());
		if (flag == 1) ba.setAlignLeft();
		if (flag == 2) ba.setAlignRight();
		ba.calculateAlignment();
		out << ba.m_band << ' ';
		for (size_t i = 0; i < ba.getCigarCount()->size(); i++) out << (*ba.getCigarCount())[i] << (*ba.getCigarType())[i];
		out << '\n';
	}
	return 0;
}

// run <reference flags>: main.cpp's flow (main.cpp:26-106) over the host classes, with the flags of
// the reference's command line that reach the hot path (Config::setConfig, Config.cpp:118-243):
// -q -s (comma-separated files) -o -k -m -i -l -a -t -z -y -f -c [freq] -v --mapping
// --singleReads --covinfo --covfilter --merge.
static int run_flags(int argc, char **argv) {
	Config cfg;
	std::vector<std::string> qfiles;
	auto split = [](const std::string &s) {
		std::vector<std::string> out;
		std::stringstream ss(s);
		std::string item;
		while (std::getline(ss, item, ',')) out.push_back(item);
		return out;
	};
	for (int i = 2; i < argc; i++) {
		const std::string a = argv[i];
		auto val = [&]() -> std::string { return i + 1 < argc ? argv[++i] : ""; };
		if (a == "-q") qfiles = split(val());
		else if (a == "-s") cfg.vSubjectFiles = split(val());
		else if (a == "-o") cfg.outputfilename = val();
		else if (a == "-k") cfg.setKmer(std::atoi(val().c_str()));
		else if (a == "-m") cfg.setMaxMismatchRatePercent(std::atof(val().c_str()));
		else if (a == "-i") cfg.setMaxIndelRatePercent(std::atof(val().c_str()));
		else if (a == "-l") cfg.setMinimumOverlap(std::atoll(val().c_str()));
		else if (a == "-a") cfg.setAffine(val() == "a");
		else if (a == "-t") cfg.numberOfThreads = (uint16_t) std::atoi(val().c_str());
		else if (a == "-z") cfg.streamChunkSize = (uint64_t) std::atoll(val().c_str());
		else if (a == "-y") cfg.queryChunkSize = (uint64_t) std::atoll(val().c_str());
		else if (a == "-f") cfg.setHashTableFactor(std::atoi(val().c_str()));
		else if (a == "-v") cfg.setSpecialVersion(true);
		else if (a == "--mapping") cfg.setMapping(true);
		else if (a == "--singleReads") cfg.bMergedReads = true;
		else if (a == "--covinfo") cfg.bCoverageStatistic = true;
		else if (a == "--covfilter") cfg.iCoverageDepthFilter = (uint32_t) std::atoi(val().c_str());
		else if (a == "--merge") cfg.setMerging(true);
		else if (a == "-c") {
			cfg.setCallConsensus(true);
			if (i + 1 < argc && argv[i + 1][0] != '-') cfg.setConsensusFrequency(std::atof(argv[++i]));
		} else {
			std::fprintf(stderr, "unknown flag %s\n", a.c_str());
			return 2;
		}
	}
	Device dev(cfg);
	QueryDataset q = qfiles.size() == 1 ? QueryDataset(dev, qfiles[0]) : QueryDataset(dev, qfiles);
	if (!q.buildDataset()) return 3;
	HashTableLongKey table(&q);
	table.createHashTables();
	table.insertQueryDataset(&q);
	SubjectReadAligner aligner(&table);
	aligner.start();
	if (cfg.opt.mapping) {
		q.getCoverageAndWriteResults(cfg.outputfilename);
		return 0;
	}
	if (cfg.opt.call_consensus) {
		q.callConsensus();
		if (cfg.bMerging) q.mergePairedReads();
		q.writeQueryOut(cfg.outputfilename);
	}
	return 0;
}

// parse <out> <file>[,<file>...]: the query loader alone (no device): every record as
// ">atoi(name)\nsequence\n", as the reference writes reads it leaves unchanged (--singleReads)
static int run_parse(int argc, char **argv) {
	if (argc < 4) return 2;
	std::vector<std::string> files;
	std::stringstream ss(argv[3]);
	std::string item;
	while (std::getline(ss, item, ',')) files.push_back(item);
	ReadFiles rf(files);
	if (!rf.ok()) return 3;
	std::ofstream out(argv[2], std::ios_base::binary | std::ios_base::trunc);
	std::string name, s;
	while (rf.next(name, s)) out << '>' << std::atoll(name.c_str()) << '\n' << s << '\n';
	return 0;
}

// table <reads.fasta> <k> <factor> [mapping]: the HashTableMethod seam over the device index:
// table size, the reference's collision counts, empty slots, and the bucket of every key of read 0
static int run_table(int argc, char **argv) {
	if (argc < 5) return 2;
	Config cfg;
	cfg.setKmer(std::atoi(argv[3]));
	cfg.setHashTableFactor(std::atoi(argv[4]));
	cfg.setCallConsensus(true);
	if (argc > 5) cfg.setMapping(true);
	Device dev(cfg);
	QueryDataset q(dev, std::string(argv[2]));
	if (!q.buildDataset()) return 3;
	HashTableLongKey t(&q);
	HashTableMethod *m = &t;
	m->createHashTables();
	m->insertQueryDataset(&q);
	uint64_t empty = 0;
	for (uint64_t i = 0; i < m->hashTableSize; i++) empty += m->isEmptyAt((int64_t) i) ? 1 : 0;
	std::printf("size %llu collisions %lld max_single %llu empty %llu bits %lld\n", (unsigned long long) m->hashTableSize,
			(long long) m->numberOfHashCollision, (unsigned long long) m->maxSingleHashCollision,
			(unsigned long long) empty, (long long) m->iBitOfReadId);
	const std::string r0 = *q.getReadFromID(0);
	for (int64_t p = 0; p + m->hashKeyLength <= (int64_t) r0.size(); p += m->hashKeyLength) {
		std::vector<int64_t> *l = m->getListOfReads(r0, p, m->hashKeyLength);
		std::printf("key %lld:", (long long) p);
		if (l) for (int64_t e : *l) std::printf(" %lld", (long long) e);
		// the same pointer again: the table keeps its lists
		std::printf(" same %d\n", (int) (l == m->getListOfReads(r0, p, m->hashKeyLength)));
	}
	return 0;
}

int main(int argc, char **argv) {
	if (argc < 2) return 2;
	const std::string mode = argv[1];
	try {
		if (mode == "correct") return run_path(true, false, argc, argv);
		if (mode == "aln") return run_path(false, false, argc, argv);
		if (mode == "mapping") return run_path(true, true, argc, argv);
		if (mode == "band") return run_band(argc, argv);
		if (mode == "run") return run_flags(argc, argv);
		if (mode == "parse") return run_parse(argc, argv);
		if (mode == "table") return run_table(argc, argv);
	} catch (const std::exception &e) {
		std::fprintf(stderr, "%s\n", e.what());
		return 1;
	}
	return 2;
}
