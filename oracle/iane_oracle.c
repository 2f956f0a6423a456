/*
 * TEST INFRASTRUCTURE — CPU oracle of the IANE convolution path.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker (or the timed CPU baseline).
 * The product (iane_amd, libiane_hip.so) never calls it.
 *
 * A plain-C restatement of the reference's CPU convolve:
 *   - tap loop in kernel storage order, axis 0 fastest
 *     (include/iane/algorithms/convolution.hpp:36-49, include/iane/math/index.hpp:53-64);
 *   - per tap sum = fma((double)sample, (double)coefficient, sum)
 *     (convolution.hpp:44-45, accumulator type concepts/element.hpp:17-23);
 *   - sample at p + c - q, c = extent/2 (convolution.hpp:38-41), border rule per axis
 *     (include/iane/locator/image_locator.hpp:40-56);
 *   - output conversion (convolution.hpp:51-75);
 *   - the cpu_policy partition: tasks = ceil(count/grain), workers = min(threads, tasks),
 *     contiguous balanced task ranges, caller is worker 0
 *     (include/iane/algorithms/for_each.hpp:75-134).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum { O_U8 = 0, O_U16 = 1, O_I32 = 2, O_F32 = 3, O_F64 = 4 };
enum { O_CLAMP = 0, O_ZERO = 1 };

typedef struct {
	char* ptr;
	int rank;
	int dtype;
	int extent[4];
	int64_t stride[4]; /* bytes */
} oracle_view;

static double load_value(const char* p, int type)
{
	switch (type) {
	case O_U8: return (double)*(const uint8_t*)p;
	case O_U16: return (double)*(const uint16_t*)p;
	case O_I32: return (double)*(const int32_t*)p;
	case O_F32: return (double)*(const float*)p;
	default: return *(const double*)p;
	}
}

/* convolution.hpp:51-75 */
static void store_value(char* p, int type, double sum)
{
	switch (type) {
	case O_U8: {
		uint8_t v;
		if (!(sum > 0.0)) {
			v = 0;
		} else if (sum >= 255.0) {
			v = 255;
		} else {
			unsigned int integer = (unsigned int)sum;
			v = (uint8_t)(integer + (sum - integer >= 0.5));
		}
		*(uint8_t*)p = v;
		break;
	}
	case O_U16: {
		uint16_t v;
		if (sum != sum) v = 0;
		else if (sum <= 0.0) v = 0;
		else if (sum >= 65535.0) v = 65535;
		else v = (uint16_t)sum;
		*(uint16_t*)p = v;
		break;
	}
	case O_I32: {
		int32_t v;
		if (sum != sum) v = 0;
		else if (sum <= -2147483648.0) v = INT32_MIN;
		else if (sum >= 2147483647.0) v = INT32_MAX;
		else v = (int32_t)sum;
		*(int32_t*)p = v;
		break;
	}
	case O_F32: *(float*)p = (float)sum; break;
	default: *(double*)p = sum; break;
	}
}

typedef struct {
	const oracle_view* src;
	const oracle_view* ker;
	const oracle_view* dst;
	int border;
	size_t count;
	size_t grain;
	size_t tasks;
	size_t workers;
	size_t worker;
} job;

/* image_locator.hpp:40-56: returns 0 when the zero rule makes the sample 0 */
static int locate(const oracle_view* v, const int* coord, int border, int64_t* offset)
{
	int64_t off = 0;
	for (int a = 0; a < v->rank; ++a) {
		int64_t c = coord[a];
		if (c < 0 || c >= v->extent[a]) {
			if (border == O_ZERO || v->extent[a] <= 0) {
				return 0;
			}
			c = c < 0 ? 0 : v->extent[a] - 1;
		}
		off += c * v->stride[a];
	}
	*offset = off;
	return 1;
}

static int advance(int* pos, const int* size, int rank)
{
	for (int i = 0; i < rank; ++i) {
		if (++pos[i] < size[i]) {
			return 1;
		}
		pos[i] = 0;
	}
	return 0;
}

static void coordinate_at(size_t index, const int* size, int rank, int* pos)
{
	for (int i = 0; i < rank - 1; ++i) {
		pos[i] = (int)(index % (size_t)size[i]);
		index /= (size_t)size[i];
	}
	pos[rank - 1] = (int)index;
}

static void filter_one(const job* j, const int* p)
{
	const oracle_view* ker = j->ker;
	const int rank = ker->rank;
	int q[4] = {0, 0, 0, 0};
	int s[4];
	double sum = 0.0;
	do {
		int64_t koff = 0;
		for (int a = 0; a < rank; ++a) {
			s[a] = p[a] + ker->extent[a] / 2 - q[a];
			koff += (int64_t)q[a] * ker->stride[a];
		}
		int64_t off;
		const double sample = locate(j->src, s, j->border, &off) ? load_value(j->src->ptr + off, j->src->dtype) : 0.0;
		sum = fma(sample, load_value(ker->ptr + koff, ker->dtype), sum);
	} while (advance(q, ker->extent, rank));
	int64_t doff = 0;
	for (int a = 0; a < rank; ++a) {
		doff += (int64_t)p[a] * j->dst->stride[a];
	}
	store_value(j->dst->ptr + doff, j->dst->dtype, sum);
}

/* for_each.hpp:100-111 */
static void* run_worker(void* arg)
{
	const job* j = (const job*)arg;
	const size_t base = j->tasks / j->workers;
	const size_t extra = j->tasks % j->workers;
	const size_t first = j->worker * base + (j->worker < extra ? j->worker : extra);
	const size_t last = first + base + (j->worker < extra);
	const int rank = j->dst->rank;
	for (size_t task = first; task < last; ++task) {
		const size_t begin = task * j->grain;
		const size_t length = j->grain < j->count - begin ? j->grain : j->count - begin;
		int pos[4] = {0, 0, 0, 0};
		coordinate_at(begin, j->dst->extent, rank, pos);
		for (size_t i = 0; i < length; ++i) {
			filter_one(j, pos);
			advance(pos, j->dst->extent, rank);
		}
	}
	return NULL;
}

/* Returns 0 ok, 1 invalid argument (same checks as convolution.hpp:107-113, transform.hpp:51-63). */
int oracle_convolve(const oracle_view* src, const oracle_view* ker, const oracle_view* dst, int border,
	size_t threads, size_t grain)
{
	if (src->rank != ker->rank || src->rank != dst->rank || src->rank < 1 || src->rank > 4) {
		return 1;
	}
	for (int a = 0; a < ker->rank; ++a) {
		if (ker->extent[a] <= 0 || ker->extent[a] % 2 == 0) {
			return 1;
		}
		if (src->extent[a] != dst->extent[a] || src->extent[a] < 0) {
			return 1;
		}
	}
	if (threads == 0 || grain == 0) {
		return 1;
	}
	size_t count = 1;
	for (int a = 0; a < dst->rank; ++a) {
		count *= (size_t)dst->extent[a];
	}
	if (count == 0) {
		return 0;
	}
	const size_t tasks = 1 + (count - 1) / grain;
	const size_t workers = threads < tasks ? threads : tasks;
	job* jobs = (job*)calloc(workers, sizeof(job));
	pthread_t* ids = (pthread_t*)calloc(workers, sizeof(pthread_t));
	for (size_t w = 0; w < workers; ++w) {
		jobs[w] = (job){src, ker, dst, border, count, grain, tasks, workers, w};
	}
	for (size_t w = 1; w < workers; ++w) {
		pthread_create(&ids[w], NULL, run_worker, &jobs[w]);
	}
	run_worker(&jobs[0]);
	for (size_t w = 1; w < workers; ++w) {
		pthread_join(ids[w], NULL);
	}
	free(ids);
	free(jobs);
	return 0;
}
