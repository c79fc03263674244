"""CPU emulation of the serving batch's work-queue protocol (no GPU).

batch_queue_kernel (csrc/batch.hip) runs a persistent grid whose workgroups
take items from one shared counter: thread 0 increments it (relaxed; the
increment for the following item is issued when an item starts), the index
is broadcast through LDS between workgroup barriers, and every thread leaves
the loop once the index is past the end.  This test compiles
the same loop with host threads standing in for the workgroups' threads
(std::atomic for the counter, std::barrier for the workgroup barrier, a
shared int for the LDS word), with random delays, and checks that every item
is taken exactly once, that every workgroup finishes, and that no workgroup
takes more than n_items + 1 indices -- the termination argument of the
kernel, independent of scheduling.
"""
import os
import subprocess

import pytest

SRC = r"""
#include <atomic>
#include <barrier>
#include <cstdio>
#include <random>
#include <thread>
#include <vector>

int main(int argc, char** argv) {
  const int G = atoi(argv[1]), T = atoi(argv[2]), n_items = atoi(argv[3]);
  std::atomic<int> head{0};
  std::vector<std::atomic<int>> taken(n_items);
  for (auto& t : taken) t = 0;
  std::vector<int> fetches(G, 0);
  std::vector<std::thread> threads;
  for (int g = 0; g < G; ++g) {
    auto* bar = new std::barrier<>(T);
    auto* next = new int(0);
    for (int t = 0; t < T; ++t) {
      threads.emplace_back([=, &head, &taken, &fetches] {
        std::mt19937 rng(977 * g + t);
        if (t == 0) {
          *next = head.fetch_add(1, std::memory_order_relaxed);
          ++fetches[g];
        }
        bar->arrive_and_wait();
        int j = *next;
        bar->arrive_and_wait();
        while (j < n_items) {
          int ahead = 0;
          if (t == 0) {  // the index of the following item, taken when this one starts
            ahead = head.fetch_add(1, std::memory_order_relaxed);
            ++fetches[g];
          }
          if (t == 0) taken[j].fetch_add(1);
          if (rng() % 4 == 0) std::this_thread::yield();  // uneven item times
          bar->arrive_and_wait();  // the item's own barriers (block_sum)
          if (t == 0) *next = ahead;
          bar->arrive_and_wait();
          j = *next;
          bar->arrive_and_wait();
        }
      });
    }
  }
  for (auto& th : threads) th.join();
  int bad = 0, most = 0;
  for (int j = 0; j < n_items; ++j) bad += taken[j] != 1;
  for (int g = 0; g < G; ++g) most = fetches[g] > most ? fetches[g] : most;
  printf("%d %d %d\n", bad, most, head.load());
  return 0;
}
"""


@pytest.fixture(scope="module")
def emulator(tmp_path_factory):
    d = tmp_path_factory.mktemp("queue")
    src = os.path.join(d, "queue.cpp")
    exe = os.path.join(d, "queue")
    with open(src, "w") as f:
        f.write(SRC)
    subprocess.run(["g++", "-O2", "-std=c++20", "-pthread", src, "-o", exe], check=True)
    return exe


@pytest.mark.parametrize("G,T,n", [(1, 4, 50), (8, 4, 3), (16, 4, 200), (32, 2, 1000), (4, 8, 0)])
def test_every_item_once_and_every_workgroup_ends(emulator, G, T, n):
    out = subprocess.run([emulator, str(G), str(T), str(n)], capture_output=True, text=True, timeout=60, check=True)
    bad, most, head = map(int, out.stdout.split())
    assert bad == 0                 # each item taken exactly once
    assert most <= n + 1            # the bound the kernel's termination rests on
    assert head == n + G            # every workgroup's last fetch is its exit
