// Fork-join loop for the host mesh work: the per-level set algebra of a mesh
// version is independent across levels.  A persistent pool of workers is
// woken per loop (spawning threads per loop costs more than the small
// per-level tasks of a mesh rebuild).
#pragma once

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdlib>
#include <exception>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace smr
{
    class HostPool
    {
      public:
        static HostPool& get()
        {
            static HostPool pool;
            return pool;
        }

        int threads() const { return nthreads_; }

        // f(i) for i in [0, n); the first exception of any f(i) is rethrown
        void run(int n, const std::function<void(int)>& f)
        {
            if (n <= 0)
            {
                return;
            }
            if (nthreads_ <= 1 || n == 1 || busy_.exchange(true))
            {
                // single-threaded, or called from inside a pool task
                for (int i = 0; i < n; ++i)
                {
                    f(i);
                }
                return;
            }
            {
                std::lock_guard<std::mutex> lk(m_);
                job_     = &f;
                n_       = n;
                next_    = 0;
                err_     = nullptr;
                pending_ = int(workers_.size());
                ++generation_;
            }
            cv_.notify_all();
            work();
            {
                std::unique_lock<std::mutex> lk(m_);
                done_cv_.wait(lk, [&] { return pending_ == 0; });
                job_ = nullptr;
            }
            busy_ = false;
            if (err_)
            {
                std::rethrow_exception(err_);
            }
        }

        ~HostPool()
        {
            {
                std::lock_guard<std::mutex> lk(m_);
                stop_ = true;
                ++generation_;
            }
            cv_.notify_all();
            for (auto& t : workers_)
            {
                t.join();
            }
        }

      private:
        HostPool()
        {
            const char* e = std::getenv("SMR_HOST_THREADS");
            int v         = e ? std::atoi(e) : int(std::thread::hardware_concurrency());
            nthreads_     = std::max(1, std::min(v, 16));
            for (int t = 1; t < nthreads_; ++t)
            {
                workers_.emplace_back([this] { loop(); });
            }
        }

        void work()
        {
            for (int i = next_++; i < n_; i = next_++)
            {
                try
                {
                    (*job_)(i);
                }
                catch (...)
                {
                    std::lock_guard<std::mutex> lk(err_m_);
                    if (!err_)
                    {
                        err_ = std::current_exception();
                    }
                }
            }
        }

        void loop()
        {
            uint64_t seen = 0;
            for (;;)
            {
                {
                    std::unique_lock<std::mutex> lk(m_);
                    cv_.wait(lk, [&] { return generation_ != seen; });
                    seen = generation_;
                    if (stop_)
                    {
                        return;
                    }
                }
                work();
                {
                    std::lock_guard<std::mutex> lk(m_);
                    if (--pending_ == 0)
                    {
                        done_cv_.notify_one();
                    }
                }
            }
        }

        int nthreads_ = 1;
        std::vector<std::thread> workers_;
        std::mutex m_, err_m_;
        std::condition_variable cv_, done_cv_;
        const std::function<void(int)>* job_ = nullptr;
        std::atomic<int> next_{0};
        int n_                = 0;
        int pending_          = 0;
        uint64_t generation_  = 0;
        bool stop_            = false;
        std::atomic<bool> busy_{false};
        std::exception_ptr err_;
    };

    template <class F>
    void parallel_for(int n, F&& f)
    {
        const std::function<void(int)> fn = f;
        HostPool::get().run(n, fn);
    }
} // namespace smr
